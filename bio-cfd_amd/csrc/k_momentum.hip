// Explicit momentum step on gfx950 (reference nsMomentum2order,
// src/NavierStokes.f90:252-912): ut = u + dt*(-convection + laplacian/re
// - grad p) per fluid cell, third-order non-uniform upwind convection away
// from the body and first-order donor-cell convection next to it.  FP64, same
// operand order as the reference.
//
// Thin-surface (TS) cells: for fluid cells face-adjacent to a TS cell the
// reference temporarily substitutes ghost values into neighbouring u,v,w,p
// entries, evaluates the stencil and then writes the "t" ghost values back
// (NavierStokes.f90:310-490, 860-907), all in place and in fluid-list order.
// That order is reproduced exactly: cells whose 5x5x5 neighbourhood contains
// a position that any TS-adjacent cell writes are processed afterwards, one
// by one in list order, by a single workgroup; every other fluid cell reads
// only untouched values and runs in the dense parallel kernel.
#include <cmath>
#include <cstring>
#include <type_traits>

#include "biocfd_rt.hpp"
#include "dual.hpp"
#include "rdiv.hpp"

namespace bcfd {

struct MomArgs {
  Dims d;
  const uint8_t *cell, *cell2;
  const double *dx, *dy, *dz;
  const double *ca[BTB_N], *ck[BTB_N];
  // FP32 copies of the upwind tables and spacings (the FP32 tendency)
  const float *ca32[BTB_N], *ck32[BTB_N];
  const float *dx32, *dy32, *dz32;
  // the upwind denominators and their reciprocals per table (see rdiv)
  const double *rr[BTB_N];
  double dt, re, rre;   // rre = 1 / re rounded once
};

struct Fields {
  const double *u, *v, *w, *p;
};

// vel*(ca1 f+2 + ca2 f+1 + ca3 f0 + ca4 f-1 + ca5 f-2)/(ca6 hA)
//   + |vel|*(ck1 f+2 + ... + ck5 f-2)/(2 ck6 hB)       (NavierStokes.f90:619-647)
// Read-only data read at a wave-uniform address (the y and z upwind table
// rows, dy, dz): a constant-address-space pointer lets the compiler use
// scalar loads.  Nothing writes these arrays while the kernels run.
typedef const double __attribute__((address_space(4))) *cdp;
template <class R>
using cptr = const R __attribute__((address_space(4))) *;

// the upwind tables and spacings in the value type R
template <class R>
__device__ inline const R *mom_ca(const MomArgs &a, int t) {
  if constexpr (std::is_same_v<R, float>) return a.ca32[t];
  else return a.ca[t];
}
template <class R>
__device__ inline const R *mom_ck(const MomArgs &a, int t) {
  if constexpr (std::is_same_v<R, float>) return a.ck32[t];
  else return a.ck[t];
}
template <class R>
__device__ inline const R *mom_h(const MomArgs &a, int axis) {
  if constexpr (std::is_same_v<R, float>) return axis == 0 ? a.dx32 : axis == 1 ? a.dy32 : a.dz32;
  else return axis == 0 ? a.dx : axis == 1 ? a.dy : a.dz;
}

// (a, c: the table rows of the stencil's column, ca/ck + 6 (col - 1))
// the type of the grid quantities and constants the stencil combines with
// operands of type T: float for the FP32 tendency, double otherwise (FP64
// values and Dual values with FP64 parts)
template <class T>
using real_of = std::conditional_t<std::is_same_v<T, float>, float, double>;

// x / y; for the FP32 tendency through the hardware reciprocal (about 1 ulp),
// correctly rounded otherwise
template <class T, class H>
__device__ inline T qdiv(T x, H y) {
  if constexpr (std::is_same_v<T, float>) return x * __builtin_amdgcn_rcpf(y);
  else return x / y;
}

// x / b from b and y = 1/b rounded once (rdiv.hpp: two FMA corrections of
// x y, the correctly rounded quotient while x, x y and the residuals are
// normal numbers).  g[0], g[1] collect the smallest and largest binary
// exponent of the numerators (0 for zero): a cell with a numerator outside
// 2^-957 <= |x| < 2^700, or with a non-finite result, is evaluated again
// with divisions.  b, y lie in [2^-40, 2^40] (checked when the tables are
// made), so x y and the residuals stay normal over that numerator range.
__device__ inline double rdiv(double x, double b, double y, int *g) {
  const int e = __builtin_amdgcn_frexp_exp(x);
  g[0] = e < g[0] ? e : g[0];
  g[1] = e > g[1] ? e : g[1];
  return rdiv_core(x, b, y);
}

// whether a cell evaluated with rdiv needs its divisions
__device__ inline bool rdiv_out_of_range(const int *g, double o1, double o2, double o3) {
  return g[0] < -956 || g[1] > 700 || !isfinite(o1) || !isfinite(o2) || !isfinite(o3);
}

// rr: the table's row {b1, 1/b1, b2, 1/b2} (accessors without rows pass
// nullptr and divide)
template <class T, class TA, class TC, class RR, class H>
__device__ inline T upwind3r(T vel, TA a, TC c, RR rr, T f2, T f1, T f0, T fm1, T fm2, H hA, H hB, int *bad) {
  if constexpr (std::is_same_v<RR, std::nullptr_t>) {
    return qdiv(vel * (a[0] * f2 + a[1] * f1 + a[2] * f0 + a[3] * fm1 + a[4] * fm2), a[5] * hA) +
           qdiv(fabs(vel) * (c[0] * f2 + c[1] * f1 + c[2] * f0 + c[3] * fm1 + c[4] * fm2), H(2) * c[5] * hB);
  } else {
    return rdiv(vel * (a[0] * f2 + a[1] * f1 + a[2] * f0 + a[3] * fm1 + a[4] * fm2), rr[0], rr[1], bad) +
           rdiv(fabs(vel) * (c[0] * f2 + c[1] * f1 + c[2] * f0 + c[3] * fm1 + c[4] * fm2), rr[2], rr[3], bad);
  }
}

// the viscous sum / re: with reciprocal rows (bad != nullptr) through rdiv
template <class T>
__device__ inline T visc_div(T x, const MomArgs &a, int *bad) {
  if constexpr (std::is_same_v<T, double>) {
    if (bad) return rdiv(x, a.re, a.rre, bad);
  }
  return qdiv(x, real_of<T>(a.re));
}


// Grid spacings around cell (i,j,k) and its linear offset, shared by the
// three component evaluations.
template <class R>
struct MomCellT {
  int64_t c, sy, sz;
  R dxr, dx, dxl, dye, dy, dyw, dzt, dz, dzb;
};
using MomCell = MomCellT<double>;

__device__ inline MomCell mom_cell(const MomArgs &a, int i, int j, int k) {
  const Dims &d = a.d;
  MomCell m;
  m.c = ix3(d, i, j, k);
  m.sy = d.NX;
  m.sz = (int64_t)d.NX * d.NY;
  m.dxr = a.dx[i]; m.dx = a.dx[i - 1]; m.dxl = a.dx[i - 2];
  m.dye = a.dy[j]; m.dy = a.dy[j - 1]; m.dyw = a.dy[j - 2];
  m.dzt = a.dz[k]; m.dz = a.dz[k - 1]; m.dzb = a.dz[k - 2];
  return m;
}

// no IB cell (cell == 2) at distance 1 and no TS cell (cell2 == 2) at
// distance 2 (check_adjacent_cell, NavierStokes.f90:1002-1022) inside the
// index windows.  The reference tests the twelve tags one after the other;
// here all twelve are loaded before any is tested (independent loads, one
// wait instead of twelve dependent round trips), which gives the same value.
// The cell2 reads two cells below are clamped to index 1: where i, j or
// k == 2 the window test fails whatever they hold.
__device__ inline bool mom_near_ok(const MomArgs &a, int i, int j, int k) {
  const Dims &d = a.d;
  const uint8_t *t1 = a.cell, *t2 = a.cell2;
  const int im = i > 2 ? i - 2 : 1, jm = j > 2 ? j - 2 : 1, km = k > 2 ? k - 2 : 1;
  const unsigned hit1 = (t1[ix3(d, i - 1, j, k)] == 2) | (t1[ix3(d, i + 1, j, k)] == 2) |
                        (t1[ix3(d, i, j - 1, k)] == 2) | (t1[ix3(d, i, j + 1, k)] == 2) |
                        (t1[ix3(d, i, j, k - 1)] == 2) | (t1[ix3(d, i, j, k + 1)] == 2);
  const unsigned hit2 = (t2[ixt(d, im, j, k)] == 2) | (t2[ixt(d, i + 2, j, k)] == 2) |
                        (t2[ixt(d, i, jm, k)] == 2) | (t2[ixt(d, i, j + 2, k)] == 2) |
                        (t2[ixt(d, i, j, km)] == 2) | (t2[ixt(d, i, j, k + 2)] == 2);
  return (i > 2) & (j > 2) & (k > 2) & ((hit1 | hit2) == 0u);
}

// x / y for a grid-derived divisor y.  P2: every axis of the block has one
// power-of-two spacing, so every grid divisor here (a spacing, the sum of
// two, or half of such a sum) is a power of two.  Then 1/y is exact and
// x * (1/y) is the same real number as x / y, rounded once: the same bits,
// subnormal results included, for about a tenth of the instructions.
template <bool P2>
__device__ inline double gdiv(double x, double y) {
  if constexpr (P2) {
    // y = 2^e with biased exponent E: 1/y has biased exponent 2046 - E
    const long long ey = __double_as_longlong(y) & 0x7FF0000000000000ll;
    return x * __longlong_as_double(0x7FE0000000000000ll - ey);
  } else {
    return x / y;
  }
}

// FP32: y = 2^e, 1/y has biased exponent 254 - E
template <bool P2>
__device__ inline float gdiv(float x, float y) {
  if constexpr (P2) {
    const int ey = __float_as_int(y) & 0x7F800000;
    return x * __int_as_float(0x7F000000 - ey);
  } else {
    return qdiv(x, y);
  }
}

// Dual (dual.hpp): the tangent-linear and adjoint passes below evaluate the
// component functions on values with a directional derivative.
template <bool P2>
__device__ inline Dual gdiv(Dual x, double y) {
  return {gdiv<P2>(x.v, y), gdiv<P2>(x.t, y)};
}

// Stencil operands straight from global memory: the cell's linear offset and
// strides, and the upwind table rows of its column in each direction.
struct GAcc {
  const double *fu, *fv, *fw, *fp;
  int64_t c, sy, sz;
  const MomArgs *a;
  int i, j, k;
  template <int di, int dj, int dk>
  __device__ double at(const double *f) const {
    return f[c + di + (int64_t)dj * sy + (int64_t)dk * sz];
  }
  template <int di, int dj, int dk> __device__ double u() const { return at<di, dj, dk>(fu); }
  template <int di, int dj, int dk> __device__ double v() const { return at<di, dj, dk>(fv); }
  template <int di, int dj, int dk> __device__ double w() const { return at<di, dj, dk>(fw); }
  template <int di, int dj, int dk> __device__ double p() const { return at<di, dj, dk>(fp); }
  // the cell's own u, v or w (C = 0, 1, 2), to which the tendency is added
  template <int C> __device__ double own() const { return at<0, 0, 0>(C == 0 ? fu : C == 1 ? fv : fw); }
  // grid spacings dx(i+o), dy(j+o), dz(k+o), o = 0, -1, -2
  template <int o> __device__ double hx() const { return a->dx[i + o]; }
  template <int o> __device__ double hy() const { return ((cdp)a->dy)[j + o]; }
  template <int o> __device__ double hz() const { return ((cdp)a->dz)[k + o]; }
  template <int T> __device__ const double *xa() const { return a->ca[T] + 6 * (i - 1); }
  template <int T> __device__ const double *xk() const { return a->ck[T] + 6 * (i - 1); }
  template <int T> __device__ cdp ya() const { return (cdp)(a->ca[T] + 6 * (j - 1)); }
  template <int T> __device__ cdp yk() const { return (cdp)(a->ck[T] + 6 * (j - 1)); }
  template <int T> __device__ cdp za() const { return (cdp)(a->ca[T] + 6 * (k - 1)); }
  template <int T> __device__ cdp zk() const { return (cdp)(a->ck[T] + 6 * (k - 1)); }
  // no reciprocal rows: the stencil divides
  template <int T> __device__ std::nullptr_t xr() const { return nullptr; }
  template <int T> __device__ std::nullptr_t yr() const { return nullptr; }
  template <int T> __device__ std::nullptr_t zr() const { return nullptr; }
  __device__ int *bad() const { return nullptr; }
};

#define U(di, dj, dk) A.template u<di, dj, dk>()
#define V(di, dj, dk) A.template v<di, dj, dk>()
#define W(di, dj, dk) A.template w<di, dj, dk>()
#define P(di, dj, dk) A.template p<di, dj, dk>()

// ut (NavierStokes.f90:600-688)
template <bool P2, class Acc, class MC>
__device__ inline auto mom_u(const MomArgs &a, const Acc &A, const MC &m, int i, int j, int k,
                               bool near_ok) {
  using T = std::decay_t<decltype(A.template u<0, 0, 0>())>;
  using R = real_of<T>;
  const R alpha = 1;
  const R dxr = m.dxr, dx = m.dx, dye = m.dye, dy = m.dy, dyw = m.dyw, dzt = m.dzt, dz = m.dz, dzb = m.dzb;
  const T u1a = U(-1, 0, 0) + U(0, 0, 0), u22 = U(-1, 0, 0) - U(0, 0, 0);
  const T u3 = U(0, 0, 0) + U(1, 0, 0), u4 = U(0, 0, 0) - U(1, 0, 0);
  const T u5 = U(0, -1, 0) + U(0, 0, 0), u6 = U(0, -1, 0) - U(0, 0, 0);
  const T u7 = U(0, 0, 0) + U(0, 1, 0), u8 = U(0, 0, 0) - U(0, 1, 0);
  const T u9 = U(0, 0, -1) + U(0, 0, 0), u10 = U(0, 0, -1) - U(0, 0, 0);
  const T u11 = U(0, 0, 0) + U(0, 0, 1), u12 = U(0, 0, 0) - U(0, 0, 1);
  const T v1a = V(0, -1, 0) + V(1, -1, 0), v22 = V(0, 0, 0) + V(1, 0, 0);
  const T w1a = W(0, 0, -1) + W(1, 0, -1), w22 = W(0, 0, 0) + W(1, 0, 0);
  const T dpdx = gdiv<P2>(P(0, 0, 0) - P(1, 0, 0), R(0.5) * (dxr + dx));
  const R dx2xr = dx + dxr, dy2ye = dy + dye, dy2yw = dy + dyw, dz2zt = dz + dzt, dz2zb = dz + dzb;
  const int nx = a.d.nx, ny = a.d.ny, nz = a.d.nz;
  T duudx, dvudy, dwudz;
  const T vu_e = V(1, 0, 0) + gdiv<P2>(dxr, dx2xr) * (V(0, 0, 0) - V(1, 0, 0));
  const T vu_w = V(1, -1, 0) + gdiv<P2>(dxr, dx2xr) * (V(0, -1, 0) - V(1, -1, 0));
  const T v_in = R(0.5) * (vu_e + vu_w);
  const T wu_n = W(1, 0, 0) + gdiv<P2>(dxr, dx2xr) * (W(0, 0, 0) - W(1, 0, 0));
  const T wu_s = W(1, 0, -1) + gdiv<P2>(dxr, dx2xr) * (W(0, 0, -1) - W(1, 0, -1));
  const T w_in = R(0.5) * (wu_n + wu_s);
  if (i > 2 && i < nx && j > 2 && j < ny + 1 && k > 2 && k < nz + 1 && near_ok) {
    const R ddy = R(0.5) * (A.template hy<-1>() + A.template hy<-2>()), ddye = R(0.5) * (A.template hy<-1>() + A.template hy<0>());
    const R ddz = R(0.5) * (A.template hz<-1>() + A.template hz<-2>()), ddzr = R(0.5) * (A.template hz<-1>() + A.template hz<0>());
    duudx = upwind3r(U(0, 0, 0), A.template xa<BTB_UU>(), A.template xk<BTB_UU>(), A.template xr<BTB_UU>(), U(2, 0, 0), U(1, 0, 0), U(0, 0, 0),
                    U(-1, 0, 0), U(-2, 0, 0), A.template hx<0>(), A.template hx<-1>(), A.bad());
    dvudy = upwind3r(v_in, A.template ya<BTB_VU>(), A.template yk<BTB_VU>(), A.template yr<BTB_VU>(), U(0, 2, 0), U(0, 1, 0), U(0, 0, 0),
                    U(0, -1, 0), U(0, -2, 0), ddye, ddy, A.bad());
    dwudz = upwind3r(w_in, A.template za<BTB_WU>(), A.template zk<BTB_WU>(), A.template zr<BTB_WU>(), U(0, 0, 2), U(0, 0, 1), U(0, 0, 0),
                    U(0, 0, -1), U(0, 0, -2), ddzr, ddz, A.bad());
  } else {
    duudx = gdiv<P2>(R(-0.25) * (u1a * u1a + alpha * fabs(u1a) * u22 - u3 * u3 - alpha * fabs(u3) * u4), A.template hx<-1>());
    dvudy = gdiv<P2>(R(-0.25) * (v1a * u5 + alpha * fabs(v1a) * u6 - v22 * u7 - alpha * fabs(v22) * u8), A.template hy<-1>());
    dwudz = gdiv<P2>(R(-0.25) * (w1a * u9 + alpha * fabs(w1a) * u10 - w22 * u11 - alpha * fabs(w22) * u12), A.template hz<-1>());
  }
  const T d2x = gdiv<P2>(R(2), dx2xr) * (gdiv<P2>(-u4, dxr) + gdiv<P2>(u22, dx));
  const T d2y = gdiv<P2>(R(2), R(0.5) * (dy2ye + dy2yw)) * (gdiv<P2>(-u8, R(0.5) * dy2ye) + gdiv<P2>(u6, R(0.5) * dy2yw));
  const T d2z = gdiv<P2>(R(2), R(0.5) * (dz2zt + dz2zb)) * (gdiv<P2>(-u12, R(0.5) * dz2zt) + gdiv<P2>(u10, R(0.5) * dz2zb));
  const T visc = visc_div(d2x + d2y + d2z, a, A.bad());
  const T res = (-duudx - dvudy - dwudz + visc);
  return A.template own<0>() + a.dt * (res + dpdx);
}

// vt (NavierStokes.f90:690-780)
template <bool P2, class Acc, class MC>
__device__ inline auto mom_v(const MomArgs &a, const Acc &A, const MC &m, int i, int j, int k,
                               bool near_ok) {
  using T = std::decay_t<decltype(A.template u<0, 0, 0>())>;
  using R = real_of<T>;
  const R alpha = 1;
  const R dxr = m.dxr, dx = m.dx, dxl = m.dxl, dye = m.dye, dy = m.dy, dzt = m.dzt, dz = m.dz, dzb = m.dzb;
  const T u13 = U(-1, 0, 0) + U(-1, 1, 0), u14 = U(0, 0, 0) + U(0, 1, 0);
  const T v3 = V(-1, 0, 0) + V(0, 0, 0), v4 = V(-1, 0, 0) - V(0, 0, 0);
  const T v5 = V(0, 0, 0) + V(1, 0, 0), v6 = V(0, 0, 0) - V(1, 0, 0);
  const T v7 = V(0, -1, 0) + V(0, 0, 0), v8 = V(0, -1, 0) - V(0, 0, 0);
  const T v9 = V(0, 0, 0) + V(0, 1, 0), v10 = V(0, 0, 0) - V(0, 1, 0);
  const T v11 = V(0, 0, -1) + V(0, 0, 0), v12 = V(0, 0, -1) - V(0, 0, 0);
  const T v13 = V(0, 0, 0) + V(0, 0, 1), v14 = V(0, 0, 0) - V(0, 0, 1);
  const T w3 = W(0, 0, -1) + W(0, 1, -1), w4 = W(0, 0, 0) + W(0, 1, 0);
  const T dpdy = gdiv<P2>(P(0, 0, 0) - P(0, 1, 0), R(0.5) * (dye + dy));
  const R dx2xr = dx + dxr, dx2xl = dx + dxl, dy2ye = dy + dye, dz2zt = dz + dzt, dz2zb = dz + dzb;
  const int nx = a.d.nx, ny = a.d.ny, nz = a.d.nz;
  const T uv_e = U(0, 1, 0) + gdiv<P2>(dye, dy2ye) * (U(0, 0, 0) - U(0, 1, 0));
  const T uv_w = U(-1, 1, 0) + gdiv<P2>(dye, dy2ye) * (U(-1, 0, 0) - U(-1, 1, 0));
  const T u_in = R(0.5) * (uv_e + uv_w);
  const T wv_n = W(0, 1, 0) + gdiv<P2>(dye, dy2ye) * (W(0, 0, 0) - W(0, 1, 0));
  const T wv_s = W(0, 1, -1) + gdiv<P2>(dye, dy2ye) * (W(0, 0, -1) - W(0, 1, -1));
  const T w_in = R(0.5) * (wv_n + wv_s);
  T duvdx, dvvdy, dwvdz;
  if (i > 2 && i < nx + 1 && j > 2 && j < ny && k > 2 && k < nz + 1 && near_ok) {
    const R ddx = R(0.5) * (A.template hx<-1>() + A.template hx<-2>()), ddxr = R(0.5) * (A.template hx<-1>() + A.template hx<0>());
    const R ddz = R(0.5) * (A.template hz<-1>() + A.template hz<-2>()), ddzr = R(0.5) * (A.template hz<-1>() + A.template hz<0>());
    duvdx = upwind3r(u_in, A.template xa<BTB_UW>(), A.template xk<BTB_UW>(), A.template xr<BTB_UW>(), V(2, 0, 0), V(1, 0, 0), V(0, 0, 0),
                    V(-1, 0, 0), V(-2, 0, 0), ddxr, ddx, A.bad());
    dvvdy = upwind3r(V(0, 0, 0), A.template ya<BTB_VV>(), A.template yk<BTB_VV>(), A.template yr<BTB_VV>(), V(0, 2, 0), V(0, 1, 0), V(0, 0, 0),
                    V(0, -1, 0), V(0, -2, 0), A.template hy<0>(), A.template hy<-1>(), A.bad());
    dwvdz = upwind3r(w_in, A.template za<BTB_WU>(), A.template zk<BTB_WU>(), A.template zr<BTB_WU>(), V(0, 0, 2), V(0, 0, 1), V(0, 0, 0),
                    V(0, 0, -1), V(0, 0, -2), ddzr, ddz, A.bad());
  } else {
    duvdx = gdiv<P2>(R(-0.25) * (u13 * v3 + alpha * fabs(u13) * v4 - u14 * v5 - alpha * fabs(u14) * v6), A.template hx<-1>());
    dvvdy = gdiv<P2>(R(-0.25) * (v7 * v7 + alpha * fabs(v7) * v8 - v9 * v9 - alpha * fabs(v9) * v10), A.template hy<-1>());
    dwvdz = gdiv<P2>(R(-0.25) * (w3 * v11 + alpha * fabs(w3) * v12 - w4 * v13 - alpha * fabs(w4) * v14), A.template hz<-1>());
  }
  const T d2x = gdiv<P2>(R(2), R(0.5) * (dx2xr + dx2xl)) * (gdiv<P2>(-v6, R(0.5) * dx2xr) + gdiv<P2>(v4, R(0.5) * dx2xl));
  const T d2y = gdiv<P2>(R(2), dy2ye) * (gdiv<P2>(-v10, dye) + gdiv<P2>(v8, dy));
  const T d2z = gdiv<P2>(R(2), R(0.5) * (dz2zt + dz2zb)) * (gdiv<P2>(-v14, R(0.5) * dz2zt) + gdiv<P2>(v12, R(0.5) * dz2zb));
  const T visc = visc_div(d2x + d2y + d2z, a, A.bad());
  const T res = (-duvdx - dvvdy - dwvdz + visc);
  return A.template own<1>() + a.dt * (res + dpdy);
}

// wt (NavierStokes.f90:782-858)
template <bool P2, class Acc, class MC>
__device__ inline auto mom_w(const MomArgs &a, const Acc &A, const MC &m, int i, int j, int k,
                               bool near_ok) {
  using T = std::decay_t<decltype(A.template u<0, 0, 0>())>;
  using R = real_of<T>;
  const R alpha = 1;
  const R dxr = m.dxr, dx = m.dx, dxl = m.dxl, dye = m.dye, dy = m.dy, dyw = m.dyw, dzt = m.dzt, dz = m.dz;
  const T u15 = U(-1, 0, 0) + U(-1, 0, 1), u16 = U(0, 0, 0) + U(0, 0, 1);
  const T v15 = V(0, -1, 0) + V(0, -1, 1), v16 = V(0, 0, 0) + V(0, 0, 1);
  const T w5 = W(-1, 0, 0) + W(0, 0, 0), w6 = W(-1, 0, 0) - W(0, 0, 0);
  const T w7 = W(0, 0, 0) + W(1, 0, 0), w8 = W(0, 0, 0) - W(1, 0, 0);
  const T w9 = W(0, -1, 0) + W(0, 0, 0), w10 = W(0, -1, 0) - W(0, 0, 0);
  const T w11 = W(0, 0, 0) + W(0, 1, 0), w12 = W(0, 0, 0) - W(0, 1, 0);
  const T w13 = W(0, 0, -1) + W(0, 0, 0), w14 = W(0, 0, -1) - W(0, 0, 0);
  const T w15 = W(0, 0, 0) + W(0, 0, 1), w16 = W(0, 0, 0) - W(0, 0, 1);
  const T dpdz = gdiv<P2>(P(0, 0, 0) - P(0, 0, 1), R(0.5) * (dzt + dz));
  const R dx2xr = dx + dxr, dx2xl = dx + dxl, dy2ye = dy + dye, dy2yw = dy + dyw, dz2zt = dz + dzt;
  const int nx = a.d.nx, ny = a.d.ny, nz = a.d.nz;
  const T uw_e = U(0, 0, 1) + gdiv<P2>(dzt, dz2zt) * (U(0, 0, 0) - U(0, 0, 1));
  const T uw_w = U(-1, 0, 1) + gdiv<P2>(dzt, dz2zt) * (U(-1, 0, 0) - U(-1, 0, 1));
  const T u_in = R(0.5) * (uw_e + uw_w);
  const T vw_n = V(0, 0, 1) + gdiv<P2>(dzt, dz2zt) * (V(0, 0, 0) - V(0, 0, 1));
  const T vw_s = V(0, -1, 1) + gdiv<P2>(dzt, dz2zt) * (V(0, -1, 0) - V(0, -1, 1));
  const T v_in = R(0.5) * (vw_n + vw_s);
  T duwdx, dvwdy, dwwdz;
  if (i > 2 && i < nx + 1 && j > 2 && j < ny + 1 && k > 2 && k < nz && near_ok) {
    const R ddx = R(0.5) * (A.template hx<-1>() + A.template hx<-2>()), ddxr = R(0.5) * (A.template hx<-1>() + A.template hx<0>());
    const R ddy = R(0.5) * (A.template hy<-1>() + A.template hy<-2>()), ddye = R(0.5) * (A.template hy<-1>() + A.template hy<0>());
    duwdx = upwind3r(u_in, A.template xa<BTB_UW>(), A.template xk<BTB_UW>(), A.template xr<BTB_UW>(), W(2, 0, 0), W(1, 0, 0), W(0, 0, 0),
                    W(-1, 0, 0), W(-2, 0, 0), ddxr, ddx, A.bad());
    dvwdy = upwind3r(v_in, A.template ya<BTB_VW>(), A.template yk<BTB_VW>(), A.template yr<BTB_VW>(), W(0, 2, 0), W(0, 1, 0), W(0, 0, 0),
                    W(0, -1, 0), W(0, -2, 0), ddye, ddy, A.bad());
    dwwdz = upwind3r(W(0, 0, 0), A.template za<BTB_WW>(), A.template zk<BTB_WW>(), A.template zr<BTB_WW>(), W(0, 0, 2), W(0, 0, 1), W(0, 0, 0),
                    W(0, 0, -1), W(0, 0, -2), A.template hz<0>(), A.template hz<-1>(), A.bad());
  } else {
    duwdx = gdiv<P2>(-R(0.5) * (u15 * w5 + alpha * fabs(u15) * w6 - u16 * w7 - alpha * fabs(u16) * w8), A.template hx<-1>());
    dvwdy = gdiv<P2>(-R(0.5) * (v15 * w9 + alpha * fabs(v15) * w10 - v16 * w11 - alpha * fabs(v16) * w12), A.template hy<-1>());
    dwwdz = gdiv<P2>(-R(0.5) * (w13 * w13 + alpha * fabs(w13) * w14 - w15 * w15 - alpha * fabs(w15) * w16), A.template hz<-1>());
  }
  const T d2x = gdiv<P2>(R(2), R(0.5) * (dx2xr + dx2xl)) * (gdiv<P2>(-w8, R(0.5) * dx2xr) + gdiv<P2>(w6, R(0.5) * dx2xl));
  const T d2y = gdiv<P2>(R(2), R(0.5) * (dy2ye + dy2yw)) * (gdiv<P2>(-w12, R(0.5) * dy2ye) + gdiv<P2>(w10, R(0.5) * dy2yw));
  const T d2z = gdiv<P2>(R(2), dz2zt) * (gdiv<P2>(-w16, dzt) + gdiv<P2>(w14, dz));
  const T visc = visc_div(d2x + d2y + d2z, a, A.bad());
  const T res = (-duwdx - dvwdy - dwwdz + visc);
  return A.template own<2>() + a.dt * (res + dpdz);
}
#undef U
#undef V
#undef W
#undef P

template <bool P2 = false>
__device__ void momentum_cell(const MomArgs &a, const Fields &F, int i, int j, int k, double &out_u,
                              double &out_v, double &out_w) {
  const MomCell m = mom_cell(a, i, j, k);
  const bool near_ok = mom_near_ok(a, i, j, k);
  const GAcc A{F.u, F.v, F.w, F.p, m.c, m.sy, m.sz, &a, i, j, k};
  out_u = mom_u<P2>(a, A, m, i, j, k, near_ok);
  out_v = mom_v<P2>(a, A, m, i, j, k, near_ok);
  out_w = mom_w<P2>(a, A, m, i, j, k, near_ok);
}

// Workgroup -> (tile along i, tile along j, plane k) over a one-dimensional
// grid of gx*gy tiles per plane times nz planes.  Consecutive workgroups go
// to the eight XCDs in turn; when a plane's tile count is a multiple of 8,
// XCD e takes the e-th eighth of every plane's tiles in row order (a band of
// j rows), so most of the j halo rows and k neighbour planes a tile reads
// were brought into that XCD's L2 by its neighbours.
__device__ inline void mom_tile(int L, int gx, int gy, int &ti, int &tj, int &tk) {
  const int T = gx * gy;
  int t;
  if ((T & 7) == 0) {
    const int per = T >> 3, l = L >> 3;
    tk = l / per;
    t = (L & 7) * per + l % per;
  } else {
    tk = L / T;
    t = L % T;
  }
  ti = t % gx;
  tj = t / gx;
}

// Dense momentum kernels over fluid cells not deferred to the ordered pass.
// A workgroup is 64 x 4 cells of one k plane, one row of i per wavefront, so
// j is wave-uniform: the y and z coefficient rows and spacings are scalar
// loads.  k_momentum evaluates the three components per cell; k_momentum_c
// one component (launched once per component), which keeps fewer values
// live per thread at the cost of re-reading the shared neighbours from the
// caches.  WAVES = 4 caps k_momentum's registers for four waves per SIMD
// (a few values spill); WAVES = 1 leaves the compiler's choice (three).  All
// give the same bits.
template <int WAVES, bool P2>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(WAVES)))
k_momentum(MomArgs a, Fields F, double *__restrict__ ut, double *__restrict__ vt, double *__restrict__ wt,
           const uint8_t *__restrict__ deferred, int gx, int gy) {
  int ti, tj, tk;
  mom_tile(blockIdx.x, gx, gy, ti, tj, tk);
  const int i = 2 + ti * 64 + threadIdx.x;
  const int j = 2 + tj * 4 + __builtin_amdgcn_readfirstlane(threadIdx.y);
  const int k = 2 + tk;
  if (i > a.d.nx + 1 || j > a.d.ny + 1) return;
  const int64_t c = ix3(a.d, i, j, k);
  const uint8_t tag = a.cell[c], dfr = deferred ? deferred[c] : 0;
  if (tag != 0 || dfr != 0) return;
  double ou, ov, ow;
  momentum_cell<P2>(a, F, i, j, k, ou, ov, ow);
  ut[c] = ou;
  vt[c] = ov;
  wt[c] = ow;
}

template <int C, bool P2>
__global__ void __launch_bounds__(256) k_momentum_c(MomArgs a, Fields F, double *__restrict__ out,
                                                    const uint8_t *__restrict__ deferred, int gx, int gy) {
  int ti, tj, tk;
  mom_tile(blockIdx.x, gx, gy, ti, tj, tk);
  const int i = 2 + ti * 64 + threadIdx.x;
  const int j = 2 + tj * 4 + __builtin_amdgcn_readfirstlane(threadIdx.y);
  const int k = 2 + tk;
  if (i > a.d.nx + 1 || j > a.d.ny + 1) return;
  const int64_t c = ix3(a.d, i, j, k);
  const uint8_t tag = a.cell[c], dfr = deferred ? deferred[c] : 0;
  if (tag != 0 || dfr != 0) return;
  const MomCell m = mom_cell(a, i, j, k);
  const bool near_ok = mom_near_ok(a, i, j, k);
  const GAcc A{F.u, F.v, F.w, F.p, m.c, m.sy, m.sz, &a, i, j, k};
  if (C == 0) out[c] = mom_u<P2>(a, A, m, i, j, k, near_ok);
  if (C == 1) out[c] = mom_v<P2>(a, A, m, i, j, k, near_ok);
  if (C == 2) out[c] = mom_w<P2>(a, A, m, i, j, k, near_ok);
}

// ---------------------------------------------------------------------------
// Plane-marching dense kernel
// ---------------------------------------------------------------------------
// A workgroup owns a 64 x TJ column tile (one row of i per wavefront) and
// marches it up KC consecutive planes.  Per plane, the u, v, w planes with a
// two-cell halo, the p plane with a one-cell halo on the high i and j sides,
// and the cell (one-cell halo) and cell2 (two-cell halo) tag planes are
// staged in LDS.  Each thread keeps its own column's u, v, w at k-2..k+2,
// p at k..k+1, cell at k-1..k+1 and cell2 at k-2..k+2 in registers and rolls
// them up a plane per step (the values two or three planes ahead are loaded
// one step early).  The x-direction upwind table rows of the tile's 64
// columns stay in LDS for the whole march.  The four operands one plane off
// and one cell aside, u(i-1,j,k+1), v(i,j-1,k+1), w(i+1,j,k-1) and
// w(i,j+1,k-1), are read from global memory.  Positions outside the arrays
// are staged as 0: the stencil reads them only where its index windows have
// already failed, as the per-plane kernels do.  Same arithmetic, same operand
// order: same bits.
template <int TJ, class R = double>
struct MarchPlane {
  static constexpr int kW = 68, kH = TJ + 4;   // two-cell halo box
  R f[3][kH * kW];                             // u, v, w of plane k
  R p[(TJ + 1) * 65];                          // p: rows j..j+TJ, columns i..i+64
  uint8_t c1[(TJ + 2) * 66];                   // cell: rows j-1..j+TJ, columns i-1..i+64
  uint8_t c2[kH * kW];                         // cell2
};

// q-th position of the halo of an H-cell box around a 64 x TJ core: rows
// 0..H-1 and H+TJ.., then the side columns of the core rows
template <int TJ, int H>
__device__ inline void halo_pos(int q, int &r, int &c) {
  constexpr int C = 64 + 2 * H;
  if (q < H * C) {
    r = q / C;
    c = q % C;
  } else if (q < 2 * H * C) {
    q -= H * C;
    r = H + TJ + q / C;
    c = q % C;
  } else {
    q -= 2 * H * C;
    r = H + q / (2 * H);
    const int s = q % (2 * H);
    c = s < H ? s : 64 + s;
  }
}

// The march's operand accessor.  R = double: the FP64 stencil; R = float:
// the FP32 tendency, whose LDS planes and tables hold FP32 values and whose
// own-column values (kept in FP64 for the final update) are rounded to FP32
// where the stencil reads them.
template <int TJ, class R = double>
struct MAcc {
  const R *su, *sv, *sw;        // own position in the staged u, v, w planes
  const R *sp;                  // own position in the staged p plane
  const R *xt;                  // own row of the staged x tables
  const MomArgs *a;
  int j, k;
  R dxr, dxc, dxl;                // dx(i), dx(i-1), dx(i-2)
  double um2, um1, u0, up1, up2, vm2, vm1, v0, vp1, vp2, wm2, wm1, w0, wp1, wp2, p0, pp1;
  double u_wt, v_st, w_eb, w_nb;   // u(i-1,j,k+1), v(i,j-1,k+1), w(i+1,j,k-1), w(i,j+1,k-1)

  template <int dk>
  __device__ static R col(double m2, double m1, double p1, double p2) {
    if constexpr (dk == -2) return R(m2);
    else if constexpr (dk == -1) return R(m1);
    else if constexpr (dk == 1) return R(p1);
    else {
      static_assert(dk == 2, "own-column offset");
      return R(p2);
    }
  }
  template <int di, int dj, int dk>
  __device__ R u() const {
    if constexpr (dk == 0) {
      if constexpr (di == 0 && dj == 0) return R(u0);
      else return su[dj * MarchPlane<TJ, R>::kW + di];
    } else if constexpr (di == 0 && dj == 0) {
      return col<dk>(um2, um1, up1, up2);
    } else {
      static_assert(di == -1 && dj == 0 && dk == 1, "u operand off the plane");
      return R(u_wt);
    }
  }
  template <int di, int dj, int dk>
  __device__ R v() const {
    if constexpr (dk == 0) {
      if constexpr (di == 0 && dj == 0) return R(v0);
      else return sv[dj * MarchPlane<TJ, R>::kW + di];
    } else if constexpr (di == 0 && dj == 0) {
      return col<dk>(vm2, vm1, vp1, vp2);
    } else {
      static_assert(di == 0 && dj == -1 && dk == 1, "v operand off the plane");
      return R(v_st);
    }
  }
  template <int di, int dj, int dk>
  __device__ R w() const {
    if constexpr (dk == 0) {
      if constexpr (di == 0 && dj == 0) return R(w0);
      else return sw[dj * MarchPlane<TJ, R>::kW + di];
    } else if constexpr (di == 0 && dj == 0) {
      return col<dk>(wm2, wm1, wp1, wp2);
    } else if constexpr (di == 1 && dj == 0 && dk == -1) {
      return R(w_eb);
    } else {
      static_assert(di == 0 && dj == 1 && dk == -1, "w operand off the plane");
      return R(w_nb);
    }
  }
  template <int C> __device__ double own() const { return C == 0 ? u0 : C == 1 ? v0 : w0; }
  template <int di, int dj, int dk>
  __device__ R p() const {
    if constexpr (di == 0 && dj == 0 && dk == 0) return R(p0);
    else if constexpr (di == 1 && dj == 0 && dk == 0) return sp[1];
    else if constexpr (di == 0 && dj == 1 && dk == 0) return sp[65];
    else {
      static_assert(di == 0 && dj == 0 && dk == 1, "p operand");
      return R(pp1);
    }
  }
  template <int o>
  __device__ R hx() const {
    if constexpr (o == 0) return dxr;
    else if constexpr (o == -1) return dxc;
    else {
      static_assert(o == -2, "x spacing");
      return dxl;
    }
  }
  template <int o> __device__ R hy() const { return ((cptr<R>)mom_h<R>(*a, 1))[j + o]; }
  template <int o> __device__ R hz() const { return ((cptr<R>)mom_h<R>(*a, 2))[k + o]; }
  template <int T>
  __device__ const R *xa() const {
    static_assert(T == BTB_UU || T == BTB_UW, "staged x table");
    return xt + (T == BTB_UU ? 0 : 2) * 384;
  }
  template <int T>
  __device__ const R *xk() const {
    static_assert(T == BTB_UU || T == BTB_UW, "staged x table");
    return xt + (T == BTB_UU ? 1 : 3) * 384;
  }
  template <int T> __device__ cptr<R> ya() const { return (cptr<R>)(mom_ca<R>(*a, T) + 6 * (j - 1)); }
  template <int T> __device__ cptr<R> yk() const { return (cptr<R>)(mom_ck<R>(*a, T) + 6 * (j - 1)); }
  template <int T> __device__ cptr<R> za() const { return (cptr<R>)(mom_ca<R>(*a, T) + 6 * (k - 1)); }
  template <int T> __device__ cptr<R> zk() const { return (cptr<R>)(mom_ck<R>(*a, T) + 6 * (k - 1)); }
  // no reciprocal rows: the stencil divides
  template <int T> __device__ std::nullptr_t xr() const { return nullptr; }
  template <int T> __device__ std::nullptr_t yr() const { return nullptr; }
  template <int T> __device__ std::nullptr_t zr() const { return nullptr; }
  __device__ int *bad() const { return nullptr; }
};

// The FP64 march accessor with the upwind denominators' reciprocal rows:
// x rows staged in LDS with the x tables, y and z rows wave-uniform; a
// numerator outside rdiv's range sets *badp and the cell is evaluated again
// through the base accessor, with divisions.
template <int TJ>
struct MAccFast : MAcc<TJ, double> {
  const double *xrt;   // own column's row of the staged x rows (UU, then UW)
  const double *yrt;   // own j row of the staged y rows (VU, VV, VW; 4 TJ apart)
  const double *zrt;   // the plane's row of the staged z rows (WU, WW; 4 KC apart)
  int ystride, zstride;
  int *badp;           // the numerators' exponent range (rdiv)
  template <int T>
  __device__ const double *xr() const {
    static_assert(T == BTB_UU || T == BTB_UW, "staged x rows");
    return xrt + (T == BTB_UU ? 0 : 256);
  }
  template <int T>
  __device__ const double *yr() const {
    static_assert(T == BTB_VU || T == BTB_VV || T == BTB_VW, "staged y rows");
    return yrt + (T == BTB_VU ? 0 : T == BTB_VV ? 1 : 2) * ystride;
  }
  template <int T>
  __device__ const double *zr() const {
    static_assert(T == BTB_WU || T == BTB_WW, "staged z rows");
    return zrt + (T == BTB_WU ? 0 : 1) * zstride;
  }
  __device__ int *bad() const { return badp; }
};

// the march of tile L (workgroup index within the block) of one block
// fix (FAST): a count, then the cells to evaluate again (n2 entries)
template <int TJ, int KC, bool P2, class R, bool FAST = false>
__device__ __forceinline__ void march_run(const MomArgs &a, const Fields &F, double *__restrict__ ut,
                                          double *__restrict__ vt, double *__restrict__ wt,
                                          const uint8_t *__restrict__ deferred, int gx, int gy, int L,
                                          int *__restrict__ fix = nullptr) {
  constexpr int NT = 64 * TJ, kW = MarchPlane<TJ>::kW;
  constexpr int NH2 = 4 * kW + 4 * TJ, NH1 = 2 * 66 + 2 * TJ, NHP = 65 + TJ, S2 = (NH2 + NT - 1) / NT;
  static_assert(NH1 <= NT && NHP <= NT, "one-cell halos: one position per thread");
  // planes alternate between two buffers, so one barrier per plane separates
  // the staging of plane k from the reads of plane k - 2
  __shared__ MarchPlane<TJ, R> S[2];
  __shared__ R tab[4][64 * 6];   // ca UU, ck UU, ca UW, ck UW rows of the tile's columns
  int ti, tj, tc;
  mom_tile(L, gx, gy, ti, tj, tc);
  const Dims &d = a.d;
  const int lane = threadIdx.x, row = __builtin_amdgcn_readfirstlane(threadIdx.y);
  const int tid = lane + 64 * row;
  const int i0 = 2 + ti * 64, j0 = 2 + tj * TJ;
  const int i = i0 + lane, j = j0 + row;
  const int kb = 2 + tc * KC, ke = kb + KC < d.nz + 2 ? kb + KC : d.nz + 2;
  const bool col_in = i <= d.NX && j <= d.NY;          // own column inside the field arrays
  const bool mine = i <= d.nx + 1 && j <= d.ny + 1;    // own column interior
  const bool col_in2 = i <= d.TX && j <= d.TY;
  const int TZ = d.nz + 3;
  const int64_t sz = (int64_t)d.NX * d.NY, szt = (int64_t)d.TX * d.TY;
  const int64_t c1 = col_in ? ix3(d, i, j, 1) : 0, t1 = col_in2 ? ixt(d, i, j, 1) : 0;
  auto ldf = [&](const double *f, int k) -> double {
    return (col_in && k >= 1 && k <= d.NZ) ? f[c1 + (int64_t)(k - 1) * sz] : 0.0;
  };
  auto ldc = [&](int k) -> uint8_t {
    return (col_in && k >= 1 && k <= d.NZ) ? a.cell[c1 + (int64_t)(k - 1) * sz] : (uint8_t)0;
  };
  auto ldc2 = [&](int k) -> uint8_t {
    return (col_in2 && k >= 1 && k <= TZ) ? a.cell2[t1 + (int64_t)(k - 1) * szt] : (uint8_t)0;
  };
  // this thread's share of a plane's halos: LDS offsets and in-plane array
  // offsets (-1: outside the array, staged as 0) fixed for the march
  int ho2[S2], ho1 = -1, hop = -1;
  int hg2[S2], ht2[S2], hg1 = -1, hgp = -1;   // in-plane offsets: below 2^31 for blocks up to 4096^2 per plane
#pragma unroll
  for (int q2 = 0; q2 < S2; ++q2) {
    const int q = tid + q2 * NT;
    ho2[q2] = -1;
    hg2[q2] = ht2[q2] = -1;
    if (q < NH2) {
      int r, cc;
      halo_pos<TJ, 2>(q, r, cc);
      const int gi = i0 - 2 + cc, gj = j0 - 2 + r;
      ho2[q2] = r * kW + cc;
      if (gi >= 1 && gi <= d.NX && gj >= 1 && gj <= d.NY) hg2[q2] = (gi - 1) + d.NX * (gj - 1);
      if (gi >= 1 && gi <= d.TX && gj >= 1 && gj <= d.TY) ht2[q2] = (gi - 1) + d.TX * (gj - 1);
    }
  }
  if (tid < NH1) {
    int r, cc;
    halo_pos<TJ, 1>(tid, r, cc);
    const int gi = i0 - 1 + cc, gj = j0 - 1 + r;
    ho1 = r * 66 + cc;
    if (gi >= 1 && gi <= d.NX && gj >= 1 && gj <= d.NY) hg1 = (gi - 1) + d.NX * (gj - 1);
  }
  if (tid < NHP) {
    const int r = tid < 65 ? TJ : tid - 65, cc = tid < 65 ? tid : 64;
    const int gi = i0 + cc, gj = j0 + r;
    hop = r * 65 + cc;
    if (gi <= d.NX && gj <= d.NY) hgp = (gi - 1) + d.NX * (gj - 1);
  }
  // the halo values of one plane, loaded one plane ahead
  double hu[S2], hv[S2], hw[S2], hp = 0.0;
  uint8_t h2[S2], h1 = 0;
  auto load_halo = [&](int k) {
    const int64_t pk = (int64_t)(k - 1) * sz, pkt = (int64_t)(k - 1) * szt;
#pragma unroll
    for (int q2 = 0; q2 < S2; ++q2) {
      hu[q2] = hv[q2] = hw[q2] = 0.0;
      h2[q2] = 0;
      if (hg2[q2] >= 0) {
        hu[q2] = F.u[hg2[q2] + pk];
        hv[q2] = F.v[hg2[q2] + pk];
        hw[q2] = F.w[hg2[q2] + pk];
      }
      if (ht2[q2] >= 0) h2[q2] = a.cell2[ht2[q2] + pkt];
    }
    h1 = hg1 >= 0 ? a.cell[hg1 + pk] : (uint8_t)0;
    hp = hgp >= 0 ? F.p[hgp + pk] : 0.0;
  };
  auto store_halo = [&](MarchPlane<TJ, R> &P) {
#pragma unroll
    for (int q2 = 0; q2 < S2; ++q2) {
      if (ho2[q2] >= 0) {
        P.f[0][ho2[q2]] = hu[q2];
        P.f[1][ho2[q2]] = hv[q2];
        P.f[2][ho2[q2]] = hw[q2];
        P.c2[ho2[q2]] = h2[q2];
      }
    }
    if (ho1 >= 0) P.c1[ho1] = h1;
    if (hop >= 0) P.p[hop] = hp;
  };

  // x-direction table rows of the tile's columns (rows 1..nx+2 exist)
  // (the table unrolled, so that each source pointer is a kernel argument,
  // not an element of a pointer array indexed at run time)
#pragma unroll
  for (int tq = 0; tq < 4; ++tq) {
    const R *src = tq == 0 ? mom_ca<R>(a, BTB_UU) : tq == 1 ? mom_ck<R>(a, BTB_UU) : tq == 2 ? mom_ca<R>(a, BTB_UW) : mom_ck<R>(a, BTB_UW);
    for (int r = tid; r < 384; r += NT) {
      const int colx = i0 + r / 6;
      R x = 0;
      if (colx <= d.nx + 2) x = src[6 * (colx - 1) + r % 6];
      tab[tq][r] = x;
    }
  }
  // with FAST, the reciprocal rows {b1, 1/b1, b2, 1/b2} of the x tables UU
  // and UW for the tile's columns, of the y tables VU, VV, VW for its rows
  // and of the z tables WU, WW for its planes (rows 1..n+2 exist)
  __shared__ double tabr[FAST ? 2 : 1][FAST ? 256 : 1];
  __shared__ double tyr[FAST ? 3 : 1][FAST ? 4 * TJ : 1];
  __shared__ double tzr[FAST ? 2 : 1][FAST ? 4 * KC : 1];
  if constexpr (FAST) {
#pragma unroll
    for (int tq = 0; tq < 2; ++tq) {
      const double *src = a.rr[tq == 0 ? BTB_UU : BTB_UW];
      for (int r = tid; r < 256; r += NT) {
        const int colx = i0 + r / 4;
        tabr[tq][r] = colx <= d.nx + 2 ? src[4 * (colx - 1) + r % 4] : 0.0;
      }
    }
#pragma unroll
    for (int tq = 0; tq < 3; ++tq) {
      const double *src = a.rr[tq == 0 ? BTB_VU : tq == 1 ? BTB_VV : BTB_VW];
      for (int r = tid; r < 4 * TJ; r += NT) {
        const int rowy = j0 + r / 4;
        tyr[tq][r] = rowy <= d.ny + 2 ? src[4 * (rowy - 1) + r % 4] : 0.0;
      }
    }
#pragma unroll
    for (int tq = 0; tq < 2; ++tq) {
      const double *src = a.rr[tq == 0 ? BTB_WU : BTB_WW];
      for (int r = tid; r < 4 * KC; r += NT) {
        const int plz = kb + r / 4;
        tzr[tq][r] = plz <= d.nz + 2 ? src[4 * (plz - 1) + r % 4] : 0.0;
      }
    }
  }
  R dxr = 0, dxc = 0, dxl = 0;
  if (mine) {
    const R *hx = mom_h<R>(a, 0);
    dxr = hx[i];
    dxc = hx[i - 1];
    dxl = hx[i - 2];
  }
  double um2 = ldf(F.u, kb - 2), um1 = ldf(F.u, kb - 1), u0 = ldf(F.u, kb), up1 = ldf(F.u, kb + 1),
         up2 = ldf(F.u, kb + 2);
  double vm2 = ldf(F.v, kb - 2), vm1 = ldf(F.v, kb - 1), v0 = ldf(F.v, kb), vp1 = ldf(F.v, kb + 1),
         vp2 = ldf(F.v, kb + 2);
  double wm2 = ldf(F.w, kb - 2), wm1 = ldf(F.w, kb - 1), w0 = ldf(F.w, kb), wp1 = ldf(F.w, kb + 1),
         wp2 = ldf(F.w, kb + 2);
  double p0 = ldf(F.p, kb), pp1 = ldf(F.p, kb + 1);
  uint8_t cm1 = ldc(kb - 1), cc0 = ldc(kb), cp1 = ldc(kb + 1);
  uint8_t c2m2 = ldc2(kb - 2), c2m1 = ldc2(kb - 1), c20 = ldc2(kb), c2p1 = ldc2(kb + 1), c2p2 = ldc2(kb + 2);
  // the deferred flag and the four operands off the plane of an interior
  // fluid cell of plane k
  uint8_t dfr = 1;
  double u_wt = 0.0, v_st = 0.0, w_eb = 0.0, w_nb = 0.0;
  auto load_off = [&](int k, uint8_t tag, uint8_t &f, double &x0, double &x1, double &x2, double &x3) {
    f = 1;
    x0 = x1 = x2 = x3 = 0.0;
    if (mine && tag == 0) {
      const int64_t c = c1 + (int64_t)(k - 1) * sz;
      f = deferred ? deferred[c] : (uint8_t)0;
      x0 = F.u[c - 1 + sz];
      x1 = F.v[c - d.NX + sz];
      x2 = F.w[c + 1 - sz];
      x3 = F.w[c + d.NX - sz];
    }
  };
  load_halo(kb);

  const int o2 = (row + 2) * kW + lane + 2, o1 = (row + 1) * 66 + lane + 1, op = row * 65 + lane;
  for (int k = kb; k < ke; ++k) {
    MarchPlane<TJ, R> &P = S[(k - kb) & 1];
    const int64_t c = c1 + (int64_t)(k - 1) * sz;
    load_off(k, cc0, dfr, u_wt, v_st, w_eb, w_nb);
    P.f[0][o2] = u0;
    P.f[1][o2] = v0;
    P.f[2][o2] = w0;
    P.c2[o2] = c20;
    P.c1[o1] = cc0;
    P.p[op] = p0;
    store_halo(P);
    __syncthreads();
    if (k + 1 < ke) load_halo(k + 1);
    if (mine && cc0 == 0 && dfr == 0) {
      MomCellT<R> m;
      m.c = c;
      m.sy = d.NX;
      m.sz = sz;
      m.dxr = dxr; m.dx = dxc; m.dxl = dxl;
      const cptr<R> hy = (cptr<R>)mom_h<R>(a, 1), hz = (cptr<R>)mom_h<R>(a, 2);
      m.dye = hy[j]; m.dy = hy[j - 1]; m.dyw = hy[j - 2];
      m.dzt = hz[k]; m.dz = hz[k - 1]; m.dzb = hz[k - 2];
      const uint8_t *s1 = &P.c1[o1], *s2 = &P.c2[o2];
      const unsigned hit1 = (s1[-1] == 2) | (s1[1] == 2) | (s1[-66] == 2) | (s1[66] == 2) | (cm1 == 2) | (cp1 == 2);
      const unsigned hit2 = (s2[-2] == 2) | (s2[2] == 2) | (s2[-2 * kW] == 2) | (s2[2 * kW] == 2) | (c2m2 == 2) |
                            (c2p2 == 2);
      const bool near_ok = (i > 2) & (j > 2) & (k > 2) & ((hit1 | hit2) == 0u);
      std::conditional_t<FAST, MAccFast<TJ>, MAcc<TJ, R>> A;
      int g[2] = {0, 0};
      if constexpr (FAST) {
        A.xrt = &tabr[0][4 * lane];
        A.yrt = &tyr[0][4 * row];
        A.zrt = &tzr[0][4 * (k - kb)];
        A.ystride = 4 * TJ;
        A.zstride = 4 * KC;
        A.badp = g;
      }
      A.su = &P.f[0][o2];
      A.sv = &P.f[1][o2];
      A.sw = &P.f[2][o2];
      A.sp = &P.p[op];
      A.xt = &tab[0][6 * lane];
      A.a = &a;
      A.j = j;
      A.k = k;
      A.dxr = dxr; A.dxc = dxc; A.dxl = dxl;
      A.um2 = um2; A.um1 = um1; A.u0 = u0; A.up1 = up1; A.up2 = up2;
      A.vm2 = vm2; A.vm1 = vm1; A.v0 = v0; A.vp1 = vp1; A.vp2 = vp2;
      A.wm2 = wm2; A.wm1 = wm1; A.w0 = w0; A.wp1 = wp1; A.wp2 = wp2;
      A.p0 = p0; A.pp1 = pp1;
      A.u_wt = u_wt; A.v_st = v_st; A.w_eb = w_eb; A.w_nb = w_nb;
      if constexpr (FAST) {
        const double ou = mom_u<P2>(a, A, m, i, j, k, near_ok);
        const double ov = mom_v<P2>(a, A, m, i, j, k, near_ok);
        const double ow = mom_w<P2>(a, A, m, i, j, k, near_ok);
        ut[c] = ou;
        vt[c] = ov;
        wt[c] = ow;
        // a numerator outside rdiv's range: the cell goes on the list that
        // k_momentum_fix evaluates again with divisions (fix[0]: the count)
        if (rdiv_out_of_range(g, ou, ov, ow)) fix[1 + atomicAdd(fix, 1)] = (int)c;
      } else {
        ut[c] = mom_u<P2>(a, A, m, i, j, k, near_ok);
        vt[c] = mom_v<P2>(a, A, m, i, j, k, near_ok);
        wt[c] = mom_w<P2>(a, A, m, i, j, k, near_ok);
      }
    }
    // roll the own column up a plane (the new values are first read two
    // planes on, so their loads have a plane's work to land)
    um2 = um1; um1 = u0; u0 = up1; up1 = up2; up2 = ldf(F.u, k + 3);
    vm2 = vm1; vm1 = v0; v0 = vp1; vp1 = vp2; vp2 = ldf(F.v, k + 3);
    wm2 = wm1; wm1 = w0; w0 = wp1; wp1 = wp2; wp2 = ldf(F.w, k + 3);
    p0 = pp1; pp1 = ldf(F.p, k + 2);
    cm1 = cc0; cc0 = cp1; cp1 = ldc(k + 2);
    c2m2 = c2m1; c2m1 = c20; c20 = c2p1; c2p1 = c2p2; c2p2 = ldc2(k + 3);
  }
}

template <int TJ, int KC, int WPE, bool P2, class R = double, bool FAST = false>
__global__ void __launch_bounds__(64 * TJ) __attribute__((amdgpu_waves_per_eu(WPE)))
k_momentum_march(MomArgs a, Fields F, double *__restrict__ ut, double *__restrict__ vt, double *__restrict__ wt,
                 const uint8_t *__restrict__ deferred, int gx, int gy, int *__restrict__ fix) {
  march_run<TJ, KC, P2, R, FAST>(a, F, ut, vt, wt, deferred, gx, gy, (int)blockIdx.x, fix);
}

// the cells of a FAST march's list again, with divisions (GAcc), by one
// workgroup (the list is normally empty), which then empties the list for
// the next call: every thread has read the count before the barrier
template <bool P2>
__global__ void __launch_bounds__(256) k_momentum_fix(MomArgs a, Fields F, double *__restrict__ ut,
                                                      double *__restrict__ vt, double *__restrict__ wt,
                                                      int *__restrict__ fix) {
  const int n = fix[0];
  __syncthreads();
  if (threadIdx.x == 0) fix[0] = 0;
  for (int q = threadIdx.x; q < n; q += blockDim.x) {
    const int c = fix[1 + q];
    const int i = 1 + c % a.d.NX, j = 1 + (c / a.d.NX) % a.d.NY, k = 1 + c / (a.d.NX * a.d.NY);
    double ou, ov, ow;
    momentum_cell<P2>(a, F, i, j, k, ou, ov, ow);
    ut[c] = ou;
    vt[c] = ov;
    wt[c] = ow;
  }
}



// The dense pass of several blocks in one launch: block b of the table owns
// workgroups first .. first + ntiles - 1 (first a multiple of 8, so that a
// workgroup's XCD is the same as in a launch of its own and the tile
// mapping keeps its XCD bands).
struct MomTask {
  MomArgs a;
  Fields F;
  double *ut, *vt, *wt;
  const uint8_t *deferred;
  int *fix;
  int gx, gy, first, ntiles;
};

template <int TJ, int KC, int WPE, bool P2, class R = double, bool FAST = false>
__global__ void __launch_bounds__(64 * TJ) __attribute__((amdgpu_waves_per_eu(WPE)))
k_momentum_march_set(const MomTask *__restrict__ tab, int ntask) {
  // (FAST: each block's own list in its task)
  int b = 0;
  while (b + 1 < ntask && (int)blockIdx.x >= tab[b + 1].first) ++b;
  const MomTask &T = tab[b];
  const int L = (int)blockIdx.x - T.first;
  if (L >= T.ntiles) return;
  march_run<TJ, KC, P2, R, FAST>(T.a, T.F, T.ut, T.vt, T.wt, T.deferred, T.gx, T.gy, L, T.fix);
}

// k_momentum_fix for the blocks of a table (one workgroup per task)
template <bool P2>
__global__ void __launch_bounds__(256) k_momentum_fix_set(const MomTask *__restrict__ tab) {
  const MomTask &T = tab[blockIdx.x];
  const int n = T.fix[0];
  __syncthreads();
  if (threadIdx.x == 0) T.fix[0] = 0;
  for (int q = threadIdx.x; q < n; q += blockDim.x) {
    const int c = T.fix[1 + q];
    const int i = 1 + c % T.a.d.NX, j = 1 + (c / T.a.d.NX) % T.a.d.NY, k = 1 + c / (T.a.d.NX * T.a.d.NY);
    double ou, ov, ow;
    momentum_cell<P2>(T.a, T.F, i, j, k, ou, ov, ow);
    T.ut[c] = ou;
    T.vt[c] = ov;
    T.wt[c] = ow;
  }
}



// ---------------------------------------------------------------------------
// TS handling
// ---------------------------------------------------------------------------
struct TsArgs {
  const int *ts_ijk;          // 3*cap (i list, j list, k list)
  int64_t cap_ts;
  const int *index_ts;        // 1-based IB ordinal
  const int *nel[BN_NNEL];
  const double *elem[BE_NELEM];
  const double *ghost[BGH_N];
  const double *x1, *y1, *z1, *xp, *yp, *zp;
  int n_ts;
};

// positions the ordered pass may write, per direction (dir 0..5 = +x,-x,+y,-y,+z,-z):
// {field (0 u,1 v,2 w,3 p), di, dj, dk, dot-test (BN_*), ghost, restore ghost}
struct Target { int8_t f, di, dj, dk, test, g, gt; };
__constant__ Target kTargets[6][6] = {
    {{0, 1, 0, 0, BN_U2, BGH_U2, BGH_U2T}, {1, 1, 0, 0, BN_V2, BGH_V2, BGH_V2T}, {1, 1, -1, 0, BN_V1, BGH_V1, BGH_V1T},
     {2, 1, 0, 0, BN_W2, BGH_W2, BGH_W2T}, {2, 1, 0, -1, BN_W1, BGH_W1, BGH_W1T}, {3, 1, 0, 0, BN_P, BGH_P, BGH_PT}},
    {{0, -2, 0, 0, BN_U1, BGH_U1, BGH_U1T}, {1, -1, 0, 0, BN_V2, BGH_V2, BGH_V2T}, {1, -1, -1, 0, BN_V1, BGH_V1, BGH_V1T},
     {2, -1, 0, 0, BN_W2, BGH_W2, BGH_W2T}, {2, -1, 0, -1, BN_W1, BGH_W1, BGH_W1T}, {3, -1, 0, 0, BN_P, BGH_P, BGH_PT}},
    {{0, 0, 1, 0, BN_U2, BGH_U2, BGH_U2T}, {0, -1, 1, 0, BN_U1, BGH_U1, BGH_U1T}, {1, 0, 1, 0, BN_V2, BGH_V2, BGH_V2T},
     {2, 0, 1, 0, BN_W2, BGH_W2, BGH_W2T}, {2, 0, 1, -1, BN_W1, BGH_W1, BGH_W1T}, {3, 0, 1, 0, BN_P, BGH_P, BGH_PT}},
    {{0, 0, -1, 0, BN_U2, BGH_U2, BGH_U2T}, {0, -1, -1, 0, BN_U1, BGH_U1, BGH_U1T}, {1, 0, -2, 0, BN_V1, BGH_V1, BGH_V1T},
     {2, 0, -1, 0, BN_W2, BGH_W2, BGH_W2T}, {2, 0, -1, -1, BN_W1, BGH_W1, BGH_W1T}, {3, 0, -1, 0, BN_P, BGH_P, BGH_PT}},
    {{0, 0, 0, 1, BN_U2, BGH_U2, BGH_U2T}, {0, -1, 0, 1, BN_U1, BGH_U1, BGH_U1T}, {1, 0, 0, 1, BN_V2, BGH_V2, BGH_V2T},
     {1, 0, -1, 1, BN_V1, BGH_V1, BGH_V1T}, {2, 0, 0, 1, BN_W2, BGH_W2, BGH_W2T}, {3, 0, 0, 1, BN_P, BGH_P, BGH_PT}},
    {{0, 0, 0, -1, BN_U2, BGH_U2, BGH_U2T}, {0, -1, 0, -1, BN_U1, BGH_U1, BGH_U1T}, {1, 0, 0, -1, BN_V2, BGH_V2, BGH_V2T},
     {1, 0, -1, -1, BN_V1, BGH_V1, BGH_V1T}, {2, 0, 0, -2, BN_W1, BGH_W1, BGH_W1T}, {3, 0, 0, -1, BN_P, BGH_P, BGH_PT}},
};
__constant__ int8_t kDir[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};

__device__ inline bool ts_adjacent(const Dims &d, const uint8_t *cell2, int i, int j, int k, unsigned *mask) {
  unsigned m = 0;
  for (int q = 0; q < 6; ++q)
    if (cell2[ixt(d, i + kDir[q][0], j + kDir[q][1], k + kDir[q][2])] == 2) m |= 1u << q;
  *mask = m;
  return m != 0;
}

// mark every position a TS-adjacent fluid cell may overwrite
__global__ void k_ts_touch(Dims d, const uint8_t *__restrict__ cell, const uint8_t *__restrict__ cell2,
                           uint8_t *__restrict__ touched) {
  const int i = 2 + blockIdx.x * blockDim.x + threadIdx.x;
  const int j = 2 + blockIdx.y * blockDim.y + threadIdx.y;
  const int k = 2 + blockIdx.z;
  if (i > d.nx + 1 || j > d.ny + 1) return;
  if (cell[ix3(d, i, j, k)] != 0) return;
  unsigned m;
  if (!ts_adjacent(d, cell2, i, j, k, &m)) return;
  for (int q = 0; q < 6; ++q) {
    if (!(m & (1u << q))) continue;
    for (int t = 0; t < 6; ++t) {
      const Target T = kTargets[q][t];
      touched[ix3(d, i + T.di, j + T.dj, k + T.dk)] = 1;
    }
  }
}

// near: every position of the arrays whose 5x5x5 neighbourhood (clipped to
// the arrays) holds a touched position; each touched position marks the
// positions around it (touched positions are few, every mark stores 1)
__global__ void k_ts_near(Dims d, const uint8_t *__restrict__ touched, uint8_t *__restrict__ near) {
  const int i = 1 + blockIdx.x * blockDim.x + threadIdx.x;
  const int j = 1 + blockIdx.y * blockDim.y + threadIdx.y;
  const int k = 1 + blockIdx.z;
  if (i > d.NX || j > d.NY) return;
  if (!touched[ix3(d, i, j, k)]) return;
  for (int dk = -2; dk <= 2; ++dk)
    for (int dj = -2; dj <= 2; ++dj)
      for (int di = -2; di <= 2; ++di) {
        const int a = i + di, b = j + dj, e = k + dk;
        if (a < 1 || b < 1 || e < 1 || a > d.NX || b > d.NY || e > d.NZ) continue;
        near[ix3(d, a, b, e)] = 1;
      }
}

// defer a fluid cell if it is TS-adjacent or its 5x5x5 neighbourhood holds a
// touched position (near)
__global__ void k_ts_defer(Dims d, const uint8_t *__restrict__ cell, const uint8_t *__restrict__ cell2,
                           const uint8_t *__restrict__ near, uint8_t *__restrict__ deferred) {
  const int i = 2 + blockIdx.x * blockDim.x + threadIdx.x;
  const int j = 2 + blockIdx.y * blockDim.y + threadIdx.y;
  const int k = 2 + blockIdx.z;
  if (i > d.nx + 1 || j > d.ny + 1) return;
  const int64_t c = ix3(d, i, j, k);
  uint8_t r = 0;
  if (cell[c] == 0) {
    unsigned m;
    r = (ts_adjacent(d, cell2, i, j, k, &m) || near[c]) ? 1 : 0;
  }
  deferred[c] = r;
}

// The ghost entry each target of a TS-adjacent deferred cell takes: the
// last TS cell (in TS order) whose element test at the target's position is
// negative, 0 for none (best[36 q + 6 dq + t] for list entry q, direction
// dq, target t).  It depends on positions and the surface only, not on the
// field values, so all cells are done at once, one workgroup per entry.  The
// workgroup scans the TS cells from the last down, blockDim at a time, and
// stops after the first chunk in which every target of the cell has a hit:
// later chunks hold smaller ordinals only.
__global__ void __launch_bounds__(256) k_ts_best(MomArgs a, const int *list, int n_list, TsArgs ts,
                                                 int *__restrict__ best) {
  __shared__ int s_best[36];
  __shared__ int s_done;
  const Dims &d = a.d;
  const int q = blockIdx.x;
  if (q >= n_list) return;
  const int lin = list[q];
  const int i = 2 + lin % d.nx;
  const int j = 2 + (lin / d.nx) % d.ny;
  const int k = 2 + lin / (d.nx * d.ny);
  unsigned m = 0;
  if (!ts_adjacent(d, a.cell2, i, j, k, &m)) return;
  if (threadIdx.x < 36) s_best[threadIdx.x] = 0;
  __syncthreads();
  const double xs[7] = {ts.x1[i], ts.x1[i - 1], ts.xp[i - 1], ts.xp[i - 1], ts.xp[i - 1], ts.xp[i - 1], ts.xp[i - 1]};
  const double ys[7] = {ts.yp[j - 1], ts.yp[j - 1], ts.y1[j], ts.y1[j - 1], ts.yp[j - 1], ts.yp[j - 1], ts.yp[j - 1]};
  const double zs[7] = {ts.zp[k - 1], ts.zp[k - 1], ts.zp[k - 1], ts.zp[k - 1], ts.z1[k], ts.z1[k - 1], ts.zp[k - 1]};
  // the seven tests in the order u2,u1,v2,v1,w2,w1,p mapped to BN_* slots
  const int test_nel[7] = {BN_U2, BN_U1, BN_V2, BN_V1, BN_W2, BN_W1, BN_P};
  for (int hi = ts.n_ts; hi >= 1; hi -= (int)blockDim.x) {
    const int nn = hi - (int)threadIdx.x;
    if (nn >= 1) {
      const int bt = ts.index_ts[nn - 1];
      bool neg[BN_NNEL];
      for (int t = 0; t < 7; ++t) {
        const int e = ts.nel[test_nel[t]][bt - 1];
        const double dn = (xs[t] - ts.elem[BE_XCENT][e - 1]) * ts.elem[BE_COSA][e - 1] +
                          (ys[t] - ts.elem[BE_YCENT][e - 1]) * ts.elem[BE_COSB][e - 1] +
                          (zs[t] - ts.elem[BE_ZCENT][e - 1]) * ts.elem[BE_COSG][e - 1];
        neg[test_nel[t]] = dn < 0;
      }
      for (int dq = 0; dq < 6; ++dq) {
        if (!(m & (1u << dq))) continue;
        for (int t = 0; t < 6; ++t)
          if (neg[kTargets[dq][t].test]) atomicMax(&s_best[dq * 6 + t], nn);
      }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      int done = 1;
      for (int slot = 0; slot < 36; ++slot)
        if (((m >> (slot / 6)) & 1u) && s_best[slot] == 0) done = 0;
      s_done = done;
    }
    __syncthreads();
    if (s_done) break;
  }
  if (threadIdx.x < 36) best[36 * (int64_t)q + threadIdx.x] = s_best[threadIdx.x];
}

// Ordered pass over the deferred cells (list in fluid order), one workgroup.
// Only a TS-adjacent cell changes u, v, w, p (its ghost substitution and the
// restore after its stencil); the cells between two TS-adjacent ones read
// the same field state and write only their own ut, vt, wt, so such a run
// is evaluated in parallel, a window of up to blockDim cells at a time, and
// each TS-adjacent cell in turn as in the reference's loop.
__global__ void __launch_bounds__(256) k_ts_ordered(MomArgs a, double *u, double *v, double *w, double *p,
                                                    double *ut, double *vt, double *wt, const int *list,
                                                    int n_list, TsArgs ts, const int *__restrict__ best) {
  __shared__ int s_first;
  const Dims &d = a.d;
  double *fld[4] = {u, v, w, p};
  auto decode = [&](int q, int &i, int &j, int &k) {
    const int lin = list[q];
    i = 2 + lin % d.nx;
    j = 2 + (lin / d.nx) % d.ny;
    k = 2 + lin / (d.nx * d.ny);
  };
  int q = 0;
  while (q < n_list) {
    // the first TS-adjacent cell in the window [q, q + blockDim)
    if (threadIdx.x == 0) s_first = n_list;
    __syncthreads();
    const int qq = q + (int)threadIdx.x;
    int ci = 0, cj = 0, ck = 0;
    unsigned cm = 0;
    if (qq < n_list) {
      decode(qq, ci, cj, ck);
      if (ts_adjacent(d, a.cell2, ci, cj, ck, &cm)) atomicMin(&s_first, qq);
    }
    __syncthreads();
    const int first = s_first;
    const int wend = q + (int)blockDim.x < n_list ? q + (int)blockDim.x : n_list;
    const int run_end = first < wend ? first : wend;
    // the run of cells that are not TS-adjacent
    if (qq < run_end) {
      Fields F{u, v, w, p};
      double ou, ov, ow;
      momentum_cell(a, F, ci, cj, ck, ou, ov, ow);
      const int64_t c = ix3(d, ci, cj, ck);
      ut[c] = ou;
      vt[c] = ov;
      wt[c] = ow;
    }
    __syncthreads();
    q = run_end;
    if (q != first || q >= n_list) continue;
    // the TS-adjacent cell q
    int i, j, k;
    decode(q, i, j, k);
    unsigned m = 0;
    const bool adj = ts_adjacent(d, a.cell2, i, j, k, &m);
    if (adj) {
      if (threadIdx.x < 36) {
        const int dq = threadIdx.x / 6, t = threadIdx.x % 6;
        const int nn = best[36 * (int64_t)q + threadIdx.x];
        if ((m & (1u << dq)) && nn > 0) {
          const Target T = kTargets[dq][t];
          fld[T.f][ix3(d, i + T.di, j + T.dj, k + T.dk)] = ts.ghost[T.g][nn - 1];
        }
      }
      __syncthreads();
    }
    if (threadIdx.x == 0) {
      Fields F{u, v, w, p};
      double ou, ov, ow;
      momentum_cell(a, F, i, j, k, ou, ov, ow);
      const int64_t c = ix3(d, i, j, k);
      ut[c] = ou;
      vt[c] = ov;
      wt[c] = ow;
      if (adj) {
        for (int dq = 0; dq < 6; ++dq) {
          if (!(m & (1u << dq))) continue;
          for (int t = 0; t < 6; ++t) {
            const Target T = kTargets[dq][t];
            fld[T.f][ix3(d, i + T.di, j + T.dj, k + T.dk)] = ts.ghost[T.gt][ts.n_ts - 1];
          }
        }
      }
    }
    __syncthreads();
    ++q;
  }
}

// ---------------------------------------------------------------------------
// The ordered pass in parallel
// ---------------------------------------------------------------------------
// In the ordered pass the only writes to u, v, w, p are a TS-adjacent cell's
// ghost substitutions at its targets (kTargets: 36 distinct positions per
// cell) and the restores that follow its own stencil; every restore of a
// target kind stores the same value, ghost[gt][n_ts - 1].  So the value a
// deferred cell q reads at a position is fixed by positions alone:
//   * q's own substitution there, if q targets it and its ghost choice
//     best[] is not 0;
//   * else the restore of the last TS-adjacent cell before q in list order
//     that targets the position (the last of its targets, in kTargets order);
//   * else the value the pass started from.
// Each deferred cell is then evaluated on its own, all at once, reading
// through that rule, and the final state of the pass is the restore of the
// last writer of every target position.  The writers of a position p are
// found from p: a cell whose target (dq, t) is p sits at p - offset(dq, t),
// so at most ten candidates per field; qmap maps a cell to its list index.

__global__ void k_ts_prep(Dims d, const uint8_t *__restrict__ cell2, const int *__restrict__ list, int n_list,
                          int *__restrict__ qmap, uint8_t *__restrict__ adjm, int clear) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n_list) return;
  const int lin = list[q];
  const int i = 2 + lin % d.nx, j = 2 + (lin / d.nx) % d.ny, k = 2 + lin / (d.nx * d.ny);
  const int64_t c = ix3(d, i, j, k);
  if (clear) {
    qmap[c] = -1;
    return;
  }
  unsigned m = 0;
  ts_adjacent(d, cell2, i, j, k, &m);
  qmap[c] = q;
  adjm[q] = (uint8_t)m;
}

struct TsView {   // passed by value: every pointer is a device address
  const int *qmap;
  const uint8_t *adjm;
  const int *best;
  const uint8_t *touched;
  const double *ghost[BGH_N];
  int n_ts;
};

// the writers of field f at cell (i,j,k): calls fn(q_w, slot) for each
template <class Fn>
__device__ inline void ts_writers(const Dims &d, const TsView &v, int f, int i, int j, int k, Fn fn) {
  for (int dq = 0; dq < 6; ++dq)
    for (int t = 0; t < 6; ++t) {
      const Target T = kTargets[dq][t];
      if (T.f != f) continue;
      const int wi = i - T.di, wj = j - T.dj, wk = k - T.dk;
      if (wi < 2 || wj < 2 || wk < 2 || wi > d.nx + 1 || wj > d.ny + 1 || wk > d.nz + 1) continue;
      const int qw = v.qmap[ix3(d, wi, wj, wk)];
      if (qw < 0 || !((v.adjm[qw] >> dq) & 1u)) continue;
      fn(qw, dq * 6 + t);
    }
}

// the 49 operands of a cell's three components: field (0 u, 1 v, 2 w, 3 p)
// and offset
struct Seed {
  signed char f, i, j, k;
};
__constant__ Seed kSeeds[49] = {
    {0, 0, 0, 0}, {0, 1, 0, 0}, {0, 2, 0, 0}, {0, -1, 0, 0}, {0, -2, 0, 0}, {0, 0, 1, 0}, {0, 0, 2, 0},
    {0, 0, -1, 0}, {0, 0, -2, 0}, {0, 0, 0, 1}, {0, 0, 0, 2}, {0, 0, 0, -1}, {0, 0, 0, -2}, {0, -1, 1, 0},
    {0, -1, 0, 1},
    {1, 0, 0, 0}, {1, 1, 0, 0}, {1, 2, 0, 0}, {1, -1, 0, 0}, {1, -2, 0, 0}, {1, 0, 1, 0}, {1, 0, 2, 0},
    {1, 0, -1, 0}, {1, 0, -2, 0}, {1, 0, 0, 1}, {1, 0, 0, 2}, {1, 0, 0, -1}, {1, 0, 0, -2}, {1, 1, -1, 0},
    {1, 0, -1, 1},
    {2, 0, 0, 0}, {2, 1, 0, 0}, {2, 2, 0, 0}, {2, -1, 0, 0}, {2, -2, 0, 0}, {2, 0, 1, 0}, {2, 0, 2, 0},
    {2, 0, -1, 0}, {2, 0, -2, 0}, {2, 0, 0, 1}, {2, 0, 0, 2}, {2, 0, 0, -1}, {2, 0, 0, -2}, {2, 1, 0, -1},
    {2, 0, 1, -1},
    {3, 0, 0, 0}, {3, 1, 0, 0}, {3, 0, 1, 0}, {3, 0, 0, 1}};

// the seed index of the operand of field F at offset (di, dj, dk)
template <int F, int di, int dj, int dk>
__device__ constexpr int sidx() {
  if constexpr (F == 3) {
    static_assert((di == 0 && dj == 0 && dk == 0) || (di == 1 && dj == 0 && dk == 0) ||
                      (di == 0 && dj == 1 && dk == 0) || (di == 0 && dj == 0 && dk == 1),
                  "p operand");
    return 45 + (di == 1 ? 1 : dj == 1 ? 2 : dk == 1 ? 3 : 0);
  } else {
    constexpr int base = 15 * F;
    if constexpr (di == 0 && dj == 0 && dk == 0) return base;
    else if constexpr (dj == 0 && dk == 0) return base + (di == 1 ? 1 : di == 2 ? 2 : di == -1 ? 3 : 4);
    else if constexpr (di == 0 && dk == 0) return base + (dj == 1 ? 5 : dj == 2 ? 6 : dj == -1 ? 7 : 8);
    else if constexpr (di == 0 && dj == 0) return base + (dk == 1 ? 9 : dk == 2 ? 10 : dk == -1 ? 11 : 12);
    else if constexpr (F == 0 && di == -1 && dj == 1 && dk == 0) return 13;
    else if constexpr (F == 0 && di == -1 && dj == 0 && dk == 1) return 14;
    else if constexpr (F == 1 && di == 1 && dj == -1 && dk == 0) return 28;
    else if constexpr (F == 1 && di == 0 && dj == -1 && dk == 1) return 29;
    else if constexpr (F == 2 && di == 1 && dj == 0 && dk == -1) return 43;
    else {
      static_assert(F == 2 && di == 0 && dj == 1 && dk == -1, "operand position");
      return 44;
    }
  }
}

// where deferred cell q at (i, j, k) reads field fi at offset (di, dj, dk)
// in the ordered pass: kind 0..3 the field itself, 4 + g the ghost array g;
// idx the entry
__device__ inline void ts_source(const Dims &d, const TsView &tv, int q, int i, int j, int k, int fi, int di,
                                 int dj, int dk, int &kind, int64_t &idx) {
  const int pi = i + di, pj = j + dj, pk = k + dk;
  const int64_t p = ix3(d, pi, pj, pk);
  kind = fi;
  idx = p;
  if (!tv.touched[p]) return;
  int last_q = -1, last_slot = -1, own_slot = -1, own_nn = 0;
  ts_writers(d, tv, fi, pi, pj, pk, [&](int qw, int slot) {
    if (qw < q) {
      if (qw > last_q || (qw == last_q && slot > last_slot)) { last_q = qw; last_slot = slot; }
    } else if (qw == q) {
      const int nn = tv.best[36 * (int64_t)q + slot];
      if (nn > 0 && slot > own_slot) {
        own_slot = slot;
        own_nn = nn;
      }
    }
  });
  if (own_slot >= 0) {
    kind = 4 + kTargets[own_slot / 6][own_slot % 6].g;
    idx = own_nn - 1;
  } else if (last_q >= 0) {
    kind = 4 + kTargets[last_slot / 6][last_slot % 6].gt;
    idx = tv.n_ts - 1;
  }
}

// The deferred cells' operands are resolved once per pass: k_ts_src finds
// where each (cell, operand) pair reads from, k_ts_gather reads the values
// into a seed-major table (entry s * n_list + q, s in kSeeds order), and the
// TS kernels take their operands from the table (one coalesced load each,
// no writer search in the stencil).  Source word: 32 * idx + kind.
static_assert(4 + BGH_N <= 32, "source word kind field");

__global__ void k_ts_src(Dims d, const int *__restrict__ list, int n_list, TsView tv, int64_t *__restrict__ src) {
  const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= 49 * (int64_t)n_list) return;
  const int s = (int)(gid / n_list), q = (int)(gid - (int64_t)s * n_list);
  const int lin = list[q];
  const int i = 2 + lin % d.nx, j = 2 + (lin / d.nx) % d.ny, k = 2 + lin / (d.nx * d.ny);
  const Seed sd = kSeeds[s];
  int kind;
  int64_t idx;
  ts_source(d, tv, q, i, j, k, sd.f, sd.i, sd.j, sd.k, kind, idx);
  src[gid] = 32 * idx + kind;
}

// the arrays a source word's kind names: u, v, w, p, then the ghost arrays
template <class T>
struct TsArrays {
  T *a[4 + BGH_N];
  __device__ T *of(int kind) const {
    T *r = a[0];
#pragma unroll
    for (int t = 1; t < 4 + BGH_N; ++t) r = kind == t ? a[t] : r;
    return r;
  }
};

__global__ void k_ts_gather(const int64_t *__restrict__ src, int64_t n, TsArrays<const double> A,
                            double *__restrict__ val) {
  const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= n) return;
  const int64_t w = src[gid];
  val[gid] = A.of((int)(w & 31))[w >> 5];
}

// stencil operands of deferred cell q as the ordered pass reads them (from
// the gathered table); spacings and upwind tables as the global accessor
struct TsAcc {
  const double *val;   // seed-major operand table
  int n;               // entries per seed
  const MomArgs *a;
  int q, i, j, k;
  template <int F, int di, int dj, int dk>
  __device__ double at() const {
    return val[(int64_t)sidx<F, di, dj, dk>() * n + q];
  }
  template <int di, int dj, int dk> __device__ double u() const { return at<0, di, dj, dk>(); }
  template <int di, int dj, int dk> __device__ double v() const { return at<1, di, dj, dk>(); }
  template <int di, int dj, int dk> __device__ double w() const { return at<2, di, dj, dk>(); }
  template <int di, int dj, int dk> __device__ double p() const { return at<3, di, dj, dk>(); }
  template <int C> __device__ double own() const { return at<C, 0, 0, 0>(); }
  template <int o> __device__ double hx() const { return a->dx[i + o]; }
  template <int o> __device__ double hy() const { return a->dy[j + o]; }
  template <int o> __device__ double hz() const { return a->dz[k + o]; }
  template <int T> __device__ const double *xa() const { return a->ca[T] + 6 * (i - 1); }
  template <int T> __device__ const double *xk() const { return a->ck[T] + 6 * (i - 1); }
  template <int T> __device__ const double *ya() const { return a->ca[T] + 6 * (j - 1); }
  template <int T> __device__ const double *yk() const { return a->ck[T] + 6 * (j - 1); }
  template <int T> __device__ const double *za() const { return a->ca[T] + 6 * (k - 1); }
  template <int T> __device__ const double *zk() const { return a->ck[T] + 6 * (k - 1); }
  // no reciprocal rows: the stencil divides
  template <int T> __device__ std::nullptr_t xr() const { return nullptr; }
  template <int T> __device__ std::nullptr_t yr() const { return nullptr; }
  template <int T> __device__ std::nullptr_t zr() const { return nullptr; }
  __device__ int *bad() const { return nullptr; }
};

// every deferred cell at once, its operands from the gathered table
__global__ void __launch_bounds__(256) k_ts_parallel(MomArgs a, const double *__restrict__ val,
                                                     double *__restrict__ ut, double *__restrict__ vt,
                                                     double *__restrict__ wt, const int *__restrict__ list,
                                                     int n_list) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n_list) return;
  const Dims &d = a.d;
  const int lin = list[q];
  const int i = 2 + lin % d.nx, j = 2 + (lin / d.nx) % d.ny, k = 2 + lin / (d.nx * d.ny);
  const MomCell m = mom_cell(a, i, j, k);
  const bool near_ok = mom_near_ok(a, i, j, k);
  const TsAcc A{val, n_list, &a, q, i, j, k};
  ut[m.c] = mom_u<false>(a, A, m, i, j, k, near_ok);
  vt[m.c] = mom_v<false>(a, A, m, i, j, k, near_ok);
  wt[m.c] = mom_w<false>(a, A, m, i, j, k, near_ok);
}

// the source words of every deferred cell's operands and their values (val:
// 49 n_list entries) for the fields u, v, w, p and the ghost arrays of the
// block's TS cells
int ts_operands(Block &b, const int *list, int64_t cnt, const TsView &tv, TsArrays<const double> fa,
                double **val) {
  IbData &ib = b.ib;
  const int64_t n = 49 * cnt;
  if (grow(&ib.ts_src, &ib.cap_ts_src, n) || grow(&ib.ts_val, &ib.cap_ts_val, 2 * n)) return -1;
  hipStream_t s = rt().stream;
  hipLaunchKernelGGL(k_ts_src, dim3(div_up(n, 256)), dim3(256), 0, s, b.d, list, (int)cnt, tv, ib.ts_src);
  hipLaunchKernelGGL(k_ts_gather, dim3(div_up(n, 256)), dim3(256), 0, s, ib.ts_src, n, fa, ib.ts_val);
  BCFD_CHECK_LAUNCH();
  *val = ib.ts_val;
  return 0;
}

// the state the ordered pass leaves: each target position holds the restore
// of its last writer
__global__ void k_ts_restore(Dims d, double *u, double *v, double *w, double *p, const int *__restrict__ list,
                             int n_list, TsView tv) {
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  const int q = gid / 36, slot = gid % 36;
  if (q >= n_list) return;
  const int dq = slot / 6;
  if (!((tv.adjm[q] >> dq) & 1u)) return;
  const int lin = list[q];
  const int i = 2 + lin % d.nx, j = 2 + (lin / d.nx) % d.ny, k = 2 + lin / (d.nx * d.ny);
  const Target T = kTargets[dq][slot % 6];
  const int pi = i + T.di, pj = j + T.dj, pk = k + T.dk;
  if (pi < 1 || pj < 1 || pk < 1 || pi > d.NX || pj > d.NY || pk > d.NZ) return;
  bool last = true;
  ts_writers(d, tv, T.f, pi, pj, pk, [&](int qw, int s2) {
    if (qw > q || (qw == q && s2 > slot)) last = false;
  });
  if (!last) return;
  double *fld[4] = {u, v, w, p};
  fld[T.f][ix3(d, pi, pj, pk)] = tv.ghost[T.gt][tv.n_ts - 1];
}

int compact_mask_lin(Block &b, const uint8_t *mask, int **out, int64_t *cap, int64_t *count);

}  // namespace bcfd

using namespace bcfd;

namespace {
// dense momentum kernel: 0 = three components per thread, 1 = one launch per
// component, 2 = as 0 with registers capped for four waves per SIMD; 3..8 =
// the plane march with 64 x TJ tiles of KC planes and registers capped for
// W waves per SIMD: (TJ, KC, W) = (4, 16, 2), (4, 32, 2), (8, 16, 2),
// (8, 32, 2), (4, 32, 3), (4, 8, 2)
constexpr int kMomVariants = 9;
// the TS pass's deferred cells: 1 = all at once (k_ts_parallel), 0 = in list
// order by one workgroup (k_ts_ordered); both give the same bits
int g_ts_parallel = 1;
int g_mom_variant = 4;   // the fastest at 256^3 on MI355X (0.72 ms against 1.02 ms for variant 0)
// grid divisions as products by exact reciprocals where the block's spacings allow
int g_mom_pow2 = 1;
// 1: the dense pass evaluates the tendency in FP32 (biocfd_momentum_precision)
int g_mom_fp32 = 0;
// 1: the march divides by the upwind denominators and re through corrected
// products with their reciprocals (biocfd_momentum_divisions)
int g_mom_rdiv = 1;

// every spacing of every axis is one and the same power of two per axis
// (the grid divisors of the momentum stencil are then exact powers of two)
bool pow2_grid(const Block &b) {
  for (int g : {BG_DX, BG_DY, BG_DZ}) {
    const std::vector<double> &h = b.hg[g];
    if (h.empty()) return false;
    int e = 0;
    const double f = std::frexp(h[0], &e);
    if (!(h[0] > 0.0) || f != 0.5 || e < -1000 || e > 1000) return false;
    for (double x : h)
      if (x != h[0]) return false;
  }
  return true;
}

// every table the momentum stencil divides by has its reciprocal rows
bool rows_ready(const MomArgs &a) {
  for (int t : {BTB_UU, BTB_VU, BTB_WU, BTB_UW, BTB_VV, BTB_VW, BTB_WW})
    if (!a.rr[t]) return false;
  return true;
}

template <int TJ, int KC, int WPE, bool P2, class R = double>
void launch_march(const MomArgs &a, const Fields &F, Block *b, hipStream_t s, const uint8_t *deferred) {
  const int gx = (int)div_up(b->d.nx, 64), gy = (int)div_up(b->d.ny, TJ), gz = (int)div_up(b->d.nz, KC);
  const dim3 g((unsigned)((int64_t)gx * gy * gz)), blk(64, TJ, 1);
  if constexpr (std::is_same_v<R, double>) {
    if (g_mom_rdiv && rows_ready(a) && b->mom_fix && b->n2 < ((int64_t)1 << 31)) {
      hipLaunchKernelGGL((k_momentum_march<TJ, KC, WPE, P2, R, true>), g, blk, 0, s, a, F, b->f[BF_UT], b->f[BF_VT],
                         b->f[BF_WT], deferred, gx, gy, b->mom_fix);
      hipLaunchKernelGGL(k_momentum_fix<P2>, dim3(1), dim3(256), 0, s, a, F, b->f[BF_UT], b->f[BF_VT],
                         b->f[BF_WT], b->mom_fix);
      return;
    }
  }
  hipLaunchKernelGGL((k_momentum_march<TJ, KC, WPE, P2, R>), g, blk, 0, s, a, F, b->f[BF_UT], b->f[BF_VT],
                     b->f[BF_WT], deferred, gx, gy, (int *)nullptr);
}

template <bool P2>
void launch_dense_p2(const MomArgs &a, const Fields &F, Block *b, hipStream_t s, const uint8_t *deferred) {
  if (g_mom_fp32) {
    // FP32 tendency of the dense pass (the deferred TS-adjacent cells stay FP64)
    launch_march<4, 32, 2, P2, float>(a, F, b, s, deferred);
    return;
  }
  switch (g_mom_variant) {
    case 3: launch_march<4, 16, 2, P2>(a, F, b, s, deferred); return;
    case 4: launch_march<4, 32, 2, P2>(a, F, b, s, deferred); return;
    case 5: launch_march<8, 16, 2, P2>(a, F, b, s, deferred); return;
    case 6: launch_march<8, 32, 2, P2>(a, F, b, s, deferred); return;
    case 7: launch_march<4, 32, 3, P2>(a, F, b, s, deferred); return;
    case 8: launch_march<4, 8, 2, P2>(a, F, b, s, deferred); return;
    default: break;
  }
  // 64 x 4 cells per workgroup, one workgroup per tile and plane
  const dim3 blk(64, 4, 1);
  const int gx = (int)div_up(b->d.nx, 64), gy = (int)div_up(b->d.ny, 4);
  const dim3 grd((unsigned)((int64_t)gx * gy * b->d.nz));
  if (g_mom_variant == 0 || g_mom_variant == 2) {
    auto kern = g_mom_variant == 0 ? k_momentum<1, P2> : k_momentum<4, P2>;
    hipLaunchKernelGGL(kern, grd, blk, 0, s, a, F, b->f[BF_UT], b->f[BF_VT], b->f[BF_WT], deferred, gx, gy);
    return;
  }
  hipLaunchKernelGGL((k_momentum_c<0, P2>), grd, blk, 0, s, a, F, b->f[BF_UT], deferred, gx, gy);
  hipLaunchKernelGGL((k_momentum_c<1, P2>), grd, blk, 0, s, a, F, b->f[BF_VT], deferred, gx, gy);
  hipLaunchKernelGGL((k_momentum_c<2, P2>), grd, blk, 0, s, a, F, b->f[BF_WT], deferred, gx, gy);
}

void launch_dense(const MomArgs &a, const Fields &F, Block *b, hipStream_t s, const uint8_t *deferred) {
  if (g_mom_pow2 && pow2_grid(*b))
    launch_dense_p2<true>(a, F, b, s, deferred);
  else
    launch_dense_p2<false>(a, F, b, s, deferred);
}
}  // namespace

extern "C" int biocfd_momentum_divisions(int reciprocal_rows) {
  BCFD_API_LOCK();
  g_mom_rdiv = reciprocal_rows != 0;
  return 0;
}

extern "C" int biocfd_momentum_precision(int fp32) {
  BCFD_API_LOCK();
  if (fp32 != 0 && fp32 != 1) {
    set_error("momentum_precision: %d is neither 0 (FP64) nor 1 (FP32 tendency)", fp32);
    return -1;
  }
  g_mom_fp32 = fp32;
  return 0;
}

extern "C" int biocfd_momentum_config(int variant, int pow2_products) {
  BCFD_API_LOCK();
  if (variant < 0 || variant >= kMomVariants) {
    set_error("momentum_config: variant %d outside 0..%d", variant, kMomVariants - 1);
    return -1;
  }
  g_mom_variant = variant;
  g_mom_pow2 = pow2_products != 0;
  return 0;
}

__global__ void k_fill_int(int *__restrict__ x, int64_t n, int value) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n) x[q] = value;
}

extern "C" int biocfd_ts_pass_config(int parallel) {
  BCFD_API_LOCK();
  g_ts_parallel = parallel != 0;
  return 0;
}

namespace {

// the block's list of cells for k_momentum_fix: a count and n2 entries,
// zero when allocated (each fix pass empties it again)
int mom_fix_ready(Block &b) {
  if (b.mom_fix && b.cap_mom_fix >= b.n2 + 1) return 0;
  if (grow(&b.mom_fix, &b.cap_mom_fix, b.n2 + 1)) return -1;
  hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, rt().stream, b.mom_fix, (int64_t)1, 0);
  BCFD_CHECK_LAUNCH();
  return 0;
}

int block_mom_args(Block *b, int id, double dt, double re, MomArgs &a) {
  for (int t = 0; t < BTB_N; ++t)
    if (!b->ca[t]) { set_error("non_uni_coeff has not been called for block %d", id); return -1; }
  if (!b->g[BG_DX]) { set_error("grids not uploaded"); return -1; }
  a.d = b->d;
  a.cell = b->tag[BT_CELL];
  a.cell2 = b->tag[BT_CELL2];
  a.dx = b->g[BG_DX]; a.dy = b->g[BG_DY]; a.dz = b->g[BG_DZ];
  for (int t = 0; t < BTB_N; ++t) {
    a.ca[t] = b->ca[t]; a.ck[t] = b->ck[t];
    a.ca32[t] = b->ca32[t]; a.ck32[t] = b->ck32[t];
  }
  a.dx32 = b->h32[0]; a.dy32 = b->h32[1]; a.dz32 = b->h32[2];
  for (int t = 0; t < BTB_N; ++t) a.rr[t] = b->rr[t];
  a.dt = dt;
  a.re = re;
  a.rre = 1.0 / re;
  return 0;
}

// before the dense pass: with TS cells, the positions the TS-adjacent cells
// may write (touched) and the cells deferred to the TS pass; *deferred =
// nullptr without TS cells
int mom_pre(Block *b, uint8_t **deferred) {
  *deferred = nullptr;
  if (b->counts[BC_TS] == 0) return 0;
  hipStream_t s = rt().stream;
  const dim3 blk(64, 4, 1), grd(div_up(b->d.nx, 64), div_up(b->d.ny, 4), (unsigned)b->d.nz);
  IbData &ib = b->ib;
  if (grow(&ib.ts_mask, &ib.cap_ts_mask, 3 * b->n2)) return -1;
  uint8_t *touched = ib.ts_mask, *dfr = ib.ts_mask + b->n2, *near = ib.ts_mask + 2 * b->n2;
  BCFD_HIP(hipMemsetAsync(ib.ts_mask, 0, 3 * b->n2, s));
  hipLaunchKernelGGL(k_ts_touch, grd, blk, 0, s, b->d, b->tag[BT_CELL], b->tag[BT_CELL2], touched);
  hipLaunchKernelGGL(k_ts_near, dim3(div_up(b->d.NX, 64), div_up(b->d.NY, 4), (unsigned)b->d.NZ), blk, 0, s, b->d,
                     touched, near);
  hipLaunchKernelGGL(k_ts_defer, grd, blk, 0, s, b->d, b->tag[BT_CELL], b->tag[BT_CELL2], near, dfr);
  BCFD_CHECK_LAUNCH();
  *deferred = dfr;
  return 0;
}

// after the dense pass: the deferred cells (TS path: deferred list, ghost
// choices, the ordered or parallel TS pass)
int mom_post(Block *b, const MomArgs &a) {
  const int n_ts = (int)b->counts[BC_TS];
  if (n_ts == 0) return 0;
  hipStream_t s = rt().stream;
  IbData &ib = b->ib;
  uint8_t *touched = ib.ts_mask, *deferred = ib.ts_mask + b->n2;
  int64_t cnt = 0;
  if (compact_mask_lin(*b, deferred, &ib.ts_list, &ib.cap_ts_list, &cnt)) return -1;
  const int *list = ib.ts_list;
  TsArgs ts{};
  ts.ts_ijk = ib.ts_ijk; ts.cap_ts = ib.cap_ts; ts.index_ts = ib.index_ts;
  for (int q = 0; q < BN_NNEL; ++q) ts.nel[q] = ib.nel[q];
  for (int q = 0; q < BE_NELEM; ++q) ts.elem[q] = ib.elem[q];
  for (int q = 0; q < BGH_N; ++q) ts.ghost[q] = ib.ghost[q];
  ts.x1 = b->g[BG_X1]; ts.y1 = b->g[BG_Y1]; ts.z1 = b->g[BG_Z1];
  ts.xp = b->g[BG_XP]; ts.yp = b->g[BG_YP]; ts.zp = b->g[BG_ZP];
  ts.n_ts = n_ts;
  if (cnt > 0) {
    if (grow(&ib.ts_best, &ib.cap_ts_best, 36 * cnt)) return -1;
    int *best = ib.ts_best;
    hipLaunchKernelGGL(k_ts_best, dim3((unsigned)cnt), dim3(256), 0, s, a, list, (int)cnt, ts, best);
    if (!g_ts_parallel) {
      hipLaunchKernelGGL(k_ts_ordered, dim3(1), dim3(256), 0, s, a, b->f[BF_U], b->f[BF_V], b->f[BF_W],
                         b->f[BF_P], b->f[BF_UT], b->f[BF_VT], b->f[BF_WT], list, (int)cnt, ts, best);
    } else {
      const int *old_map = ib.ts_qmap;
      if (grow(&ib.ts_qmap, &ib.cap_ts_qmap, b->n2) || grow(&ib.ts_adjm, &ib.cap_ts_adjm, cnt)) return -1;
      if (ib.ts_qmap != old_map)  // a new map: no cell mapped
        hipLaunchKernelGGL(k_fill_int, dim3(div_up(ib.cap_ts_qmap, 256)), dim3(256), 0, s, ib.ts_qmap,
                           ib.cap_ts_qmap, -1);
      TsView tv{};
      tv.qmap = ib.ts_qmap;
      tv.adjm = ib.ts_adjm;
      tv.best = best;
      tv.touched = touched;
      for (int g = 0; g < BGH_N; ++g) tv.ghost[g] = ts.ghost[g];
      tv.n_ts = n_ts;
      const dim3 g1(div_up(cnt, 256)), g36(div_up(36 * cnt, 256));
      hipLaunchKernelGGL(k_ts_prep, g1, dim3(256), 0, s, b->d, b->tag[BT_CELL2], list, (int)cnt, ib.ts_qmap,
                         ib.ts_adjm, 0);
      TsArrays<const double> fa{{b->f[BF_U], b->f[BF_V], b->f[BF_W], b->f[BF_P]}};
      for (int g = 0; g < BGH_N; ++g) fa.a[4 + g] = ib.ghost[g];
      double *val = nullptr;
      if (ts_operands(*b, list, cnt, tv, fa, &val)) return -1;
      hipLaunchKernelGGL(k_ts_parallel, g1, dim3(256), 0, s, a, val, b->f[BF_UT], b->f[BF_VT], b->f[BF_WT], list,
                         (int)cnt);
      hipLaunchKernelGGL(k_ts_restore, g36, dim3(256), 0, s, b->d, b->f[BF_U], b->f[BF_V], b->f[BF_W], b->f[BF_P],
                         list, (int)cnt, tv);
      hipLaunchKernelGGL(k_ts_prep, g1, dim3(256), 0, s, b->d, b->tag[BT_CELL2], list, (int)cnt, ib.ts_qmap,
                         ib.ts_adjm, 1);
    }
  }
  BCFD_CHECK_LAUNCH();
  return 0;
}


// the dense pass of blocks through one launch per power-of-two class (the
// default march shape, FP64 or the FP32 tendency); other variants block by
// block
struct MomSetTable {
  std::vector<MomTask> h;
  MomTask *d = nullptr;
  int64_t cap = 0;
};
MomSetTable g_mom_set[2];

template <bool P2>
int launch_dense_set(std::vector<MomTask> &v, MomSetTable &tt) {
  if (v.empty()) return 0;
  hipStream_t s = rt().stream;
  if (v.size() != tt.h.size() || memcmp(v.data(), tt.h.data(), sizeof(MomTask) * v.size())) {
    // a launch queued earlier may still read the old table
    BCFD_HIP(hipStreamSynchronize(s));
    if (grow(&tt.d, &tt.cap, (int64_t)v.size())) return -1;
    tt.h = v;
    BCFD_HIP(hipMemcpyAsync(tt.d, tt.h.data(), sizeof(MomTask) * v.size(), hipMemcpyHostToDevice, s));
    BCFD_HIP(hipStreamSynchronize(s));
  }
  const int64_t wgs = v.back().first + v.back().ntiles;
  bool fast = !g_mom_fp32 && g_mom_rdiv;
  for (const MomTask &t : v)
    fast = fast && rows_ready(t.a) && t.fix && (int64_t)t.a.d.NX * t.a.d.NY * t.a.d.NZ < ((int64_t)1 << 31);
  if (fast) {
    hipLaunchKernelGGL((k_momentum_march_set<4, 32, 2, P2, double, true>), dim3((unsigned)wgs), dim3(64, 4, 1), 0, s,
                       tt.d, (int)v.size());
    hipLaunchKernelGGL(k_momentum_fix_set<P2>, dim3((unsigned)v.size()), dim3(256), 0, s, (const MomTask *)tt.d);
  } else if (g_mom_fp32)
    hipLaunchKernelGGL((k_momentum_march_set<4, 32, 2, P2, float>), dim3((unsigned)wgs), dim3(64, 4, 1), 0, s, tt.d,
                       (int)v.size());
  else
    hipLaunchKernelGGL((k_momentum_march_set<4, 32, 2, P2, double>), dim3((unsigned)wgs), dim3(64, 4, 1), 0, s,
                       tt.d, (int)v.size());
  BCFD_CHECK_LAUNCH();
  return 0;
}

}  // namespace

extern "C" int biocfd_ns_momentum(int id, double dt, double re) {
  BCFD_API_LOCK();
  Block *b = get_block(id);
  if (!b) return -1;
  MomArgs a;
  if (block_mom_args(b, id, dt, re, a)) return -1;
  uint8_t *deferred = nullptr;
  if (mom_pre(b, &deferred)) return -1;
  if (mom_fix_ready(*b)) return -1;
  Fields F{b->f[BF_U], b->f[BF_V], b->f[BF_W], b->f[BF_P]};
  launch_dense(a, F, b, rt().stream, deferred);
  BCFD_CHECK_LAUNCH();
  return mom_post(b, a);
}

extern "C" int biocfd_ns_momentum_set(int first, int last, double dt, double re) {
  BCFD_API_LOCK();
  if (first < 1 || last < first || last > kMaxBlocks) {
    set_error("ns_momentum_set: bad block range %d..%d", first, last);
    return -1;
  }
  const int n = last - first + 1;
  std::vector<MomArgs> args((size_t)n);
  std::vector<uint8_t *> dfr((size_t)n);
  for (int g = first; g <= last; ++g) {
    Block *b = get_block(g);
    if (!b || block_mom_args(b, g, dt, re, args[(size_t)(g - first)])) return -1;
  }
  for (int g = first; g <= last; ++g)
    if (mom_pre(&rt().blk[g], &dfr[(size_t)(g - first)])) return -1;
  // one block, or another momentum variant: its own launch per block (the
  // table kernel reads the block's arguments from memory instead of the
  // kernel arguments: 170 against 150 us on one 96^3 block)
  if (n == 1 || (g_mom_variant != 4 && !g_mom_fp32)) {
    for (int g = first; g <= last; ++g) {
      Block *b = &rt().blk[g];
      if (mom_fix_ready(*b)) return -1;
      Fields F{b->f[BF_U], b->f[BF_V], b->f[BF_W], b->f[BF_P]};
      launch_dense(args[(size_t)(g - first)], F, b, rt().stream, dfr[(size_t)(g - first)]);
    }
    BCFD_CHECK_LAUNCH();
  } else {
    std::vector<MomTask> v[2];   // power-of-two class 0 / 1
    int64_t wgs[2] = {0, 0};
    for (int g = first; g <= last; ++g) {
      Block *b = &rt().blk[g];
      const int cls = (g_mom_pow2 && pow2_grid(*b)) ? 1 : 0;
      MomTask t;
      memset(&t, 0, sizeof t);
      t.a = args[(size_t)(g - first)];
      t.F = Fields{b->f[BF_U], b->f[BF_V], b->f[BF_W], b->f[BF_P]};
      t.ut = b->f[BF_UT];
      t.vt = b->f[BF_VT];
      t.wt = b->f[BF_WT];
      t.deferred = dfr[(size_t)(g - first)];
      if (mom_fix_ready(*b)) return -1;
      t.fix = b->mom_fix;
      t.gx = (int)div_up(b->d.nx, 64);
      t.gy = (int)div_up(b->d.ny, 4);
      t.ntiles = t.gx * t.gy * (int)div_up(b->d.nz, 32);
      t.first = (int)wgs[cls];
      wgs[cls] += (t.ntiles + 7) / 8 * 8;
      v[cls].push_back(t);
    }
    if (launch_dense_set<false>(v[0], g_mom_set[0]) || launch_dense_set<true>(v[1], g_mom_set[1])) return -1;
  }
  for (int g = first; g <= last; ++g)
    if (mom_post(&rt().blk[g], args[(size_t)(g - first)])) return -1;
  return 0;
}

// ---------------------------------------------------------------------------
// Tangent-linear and adjoint of the dense pass (the adjoint time step; no
// counterpart in the reference).  Tags, spacings and tables are the
// forward's; the operands are the forward fields u, v, w, p of the block,
// which the derivatives are taken at.  The component functions above are
// evaluated on Dual values: the tangent-linear pass carries the tangents of
// u, v, w, p (b.ad.s) into those of ut, vt, wt.  The adjoint pass is the
// hand-written reverse of the component functions (mom_u_adj, ...): the
// adjoint of every operand position the stencil reads (49 of them) is
// accumulated in registers and added once.
// ---------------------------------------------------------------------------
namespace bcfd {
namespace {

// operands with tangents read from the sensitivity arrays
struct DAcc : GAcc {
  const double *tu, *tv, *tw, *tp;
  template <int di, int dj, int dk>
  __device__ Dual u() const { return {this->template at<di, dj, dk>(fu), this->template at<di, dj, dk>(tu)}; }
  template <int di, int dj, int dk>
  __device__ Dual v() const { return {this->template at<di, dj, dk>(fv), this->template at<di, dj, dk>(tv)}; }
  template <int di, int dj, int dk>
  __device__ Dual w() const { return {this->template at<di, dj, dk>(fw), this->template at<di, dj, dk>(tw)}; }
  template <int di, int dj, int dk>
  __device__ Dual p() const { return {this->template at<di, dj, dk>(fp), this->template at<di, dj, dk>(tp)}; }
  template <int C> __device__ Dual own() const {
    if constexpr (C == 0) return u<0, 0, 0>();
    else if constexpr (C == 1) return v<0, 0, 0>();
    else return w<0, 0, 0>();
  }
};

// every operand position of the three component functions: 15 of u, v and
// w each (the own component's 13-point upwind cross and the two positions the
// other components' interpolated velocities add), 4 of p
// ---- hand-written reverse of the component functions ----
// bar[s] accumulates d(lambda . (ut, vt, wt)) / d(operand at seed s)
// (kSeeds order); the reverse follows mom_u / mom_v / mom_w statement by
// statement.  |x| has the derivative sg(x), 0 at 0 (as Dual's fabs).
__device__ inline double sg(double x) { return x > 0.0 ? 1.0 : (x < 0.0 ? -1.0 : 0.0); }

#define BU(di, dj, dk) bar[sidx<0, di, dj, dk>()]
#define BV(di, dj, dk) bar[sidx<1, di, dj, dk>()]
#define BW(di, dj, dk) bar[sidx<2, di, dj, dk>()]
#define BP(di, dj, dk) bar[sidx<3, di, dj, dk>()]
#define U(di, dj, dk) A.template u<di, dj, dk>()
#define V(di, dj, dk) A.template v<di, dj, dk>()
#define W(di, dj, dk) A.template w<di, dj, dk>()
#define P(di, dj, dk) A.template p<di, dj, dk>()

// reverse of upwind3r: b = adjoint of its value; returns the velocity's
// adjoint, f's adjoints into bf (f2, f1, f0, f-1, f-2)
template <class TA, class TC>
__device__ inline double upwind3r_adj(double vel, TA a, TC c, double f2, double f1, double f0, double fm1,
                                      double fm2, double hA, double hB, double b, double bf[5]) {
  const double sa = a[0] * f2 + a[1] * f1 + a[2] * f0 + a[3] * fm1 + a[4] * fm2;
  const double sk = c[0] * f2 + c[1] * f1 + c[2] * f0 + c[3] * fm1 + c[4] * fm2;
  const double qa = 1.0 / (a[5] * hA), qk = 1.0 / (2.0 * c[5] * hB);
  const double wa = b * vel * qa, wk = b * fabs(vel) * qk;
  for (int m = 0; m < 5; ++m) bf[m] = wa * a[m] + wk * c[m];
  return b * (sa * qa + sg(vel) * sk * qk);
}

// reverse of k (p q + |p| r - s t - |s| w): K = k * adjoint
__device__ inline void donor_adj(double K, double p, double q, double r, double s, double t, double w, double &bp,
                                 double &bq, double &br, double &bs, double &bt, double &bw) {
  bp += K * (q + sg(p) * r);
  bq += K * p;
  br += K * fabs(p);
  bs -= K * (t + sg(s) * w);
  bt -= K * s;
  bw -= K * fabs(s);
}

template <bool P2, class Acc>
__device__ inline void mom_u_adj(const MomArgs &a, const Acc &A, const MomCell &m, int i, int j, int k,
                                 bool near_ok, double lam, double bar[49]) {
  const double dxr = m.dxr, dx = m.dx, dye = m.dye, dy = m.dy, dyw = m.dyw, dzt = m.dzt, dz = m.dz, dzb = m.dzb;
  const double dx2xr = dx + dxr, dy2ye = dy + dye, dy2yw = dy + dyw, dz2zt = dz + dzt, dz2zb = dz + dzb;
  const int nx = a.d.nx, ny = a.d.ny, nz = a.d.nz;
  const double g = a.dt * lam;
  BU(0, 0, 0) += lam;
  // dpdx
  const double qp = g / (0.5 * (dxr + dx));
  BP(0, 0, 0) += qp;
  BP(1, 0, 0) -= qp;
  // viscous terms: d2x = A (-u4/dxr + u22/dx), ...
  const double bv = g / a.re;
  double bu4 = -gdiv<P2>(2.0, dx2xr) / dxr * bv, bu22 = gdiv<P2>(2.0, dx2xr) / dx * bv;
  const double ay = gdiv<P2>(2.0, 0.5 * (dy2ye + dy2yw)), az = gdiv<P2>(2.0, 0.5 * (dz2zt + dz2zb));
  double bu8 = -ay / (0.5 * dy2ye) * bv, bu6 = ay / (0.5 * dy2yw) * bv;
  double bu12 = -az / (0.5 * dz2zt) * bv, bu10 = az / (0.5 * dz2zb) * bv;
  double bu1a = 0, bu3 = 0, bu5 = 0, bu7 = 0, bu9 = 0, bu11 = 0, bv1a = 0, bv22 = 0, bw1a = 0, bw22 = 0;
  double bvin = 0, bwin = 0;
  const double rx = gdiv<P2>(dxr, dx2xr);
  if (i > 2 && i < nx && j > 2 && j < ny + 1 && k > 2 && k < nz + 1 && near_ok) {
    const double ddy = 0.5 * (A.template hy<-1>() + A.template hy<-2>()), ddye = 0.5 * (A.template hy<-1>() + A.template hy<0>());
    const double ddz = 0.5 * (A.template hz<-1>() + A.template hz<-2>()), ddzr = 0.5 * (A.template hz<-1>() + A.template hz<0>());
    const double vu_e = V(1, 0, 0) + rx * (V(0, 0, 0) - V(1, 0, 0));
    const double vu_w = V(1, -1, 0) + rx * (V(0, -1, 0) - V(1, -1, 0));
    const double v_in = 0.5 * (vu_e + vu_w);
    const double wu_n = W(1, 0, 0) + rx * (W(0, 0, 0) - W(1, 0, 0));
    const double wu_s = W(1, 0, -1) + rx * (W(0, 0, -1) - W(1, 0, -1));
    const double w_in = 0.5 * (wu_n + wu_s);
    double bf[5];
    BU(0, 0, 0) += upwind3r_adj(U(0, 0, 0), A.template xa<BTB_UU>(), A.template xk<BTB_UU>(), U(2, 0, 0), U(1, 0, 0),
                                U(0, 0, 0), U(-1, 0, 0), U(-2, 0, 0), A.template hx<0>(), A.template hx<-1>(), -g, bf);
    BU(2, 0, 0) += bf[0]; BU(1, 0, 0) += bf[1]; BU(0, 0, 0) += bf[2]; BU(-1, 0, 0) += bf[3]; BU(-2, 0, 0) += bf[4];
    bvin = upwind3r_adj(v_in, A.template ya<BTB_VU>(), A.template yk<BTB_VU>(), U(0, 2, 0), U(0, 1, 0), U(0, 0, 0),
                        U(0, -1, 0), U(0, -2, 0), ddye, ddy, -g, bf);
    BU(0, 2, 0) += bf[0]; BU(0, 1, 0) += bf[1]; BU(0, 0, 0) += bf[2]; BU(0, -1, 0) += bf[3]; BU(0, -2, 0) += bf[4];
    bwin = upwind3r_adj(w_in, A.template za<BTB_WU>(), A.template zk<BTB_WU>(), U(0, 0, 2), U(0, 0, 1), U(0, 0, 0),
                        U(0, 0, -1), U(0, 0, -2), ddzr, ddz, -g, bf);
    BU(0, 0, 2) += bf[0]; BU(0, 0, 1) += bf[1]; BU(0, 0, 0) += bf[2]; BU(0, 0, -1) += bf[3]; BU(0, 0, -2) += bf[4];
  } else {
    const double u1a = U(-1, 0, 0) + U(0, 0, 0), u22 = U(-1, 0, 0) - U(0, 0, 0);
    const double u3 = U(0, 0, 0) + U(1, 0, 0), u4 = U(0, 0, 0) - U(1, 0, 0);
    const double u5 = U(0, -1, 0) + U(0, 0, 0), u6 = U(0, -1, 0) - U(0, 0, 0);
    const double u7 = U(0, 0, 0) + U(0, 1, 0), u8 = U(0, 0, 0) - U(0, 1, 0);
    const double u9 = U(0, 0, -1) + U(0, 0, 0), u10 = U(0, 0, -1) - U(0, 0, 0);
    const double u11 = U(0, 0, 0) + U(0, 0, 1), u12 = U(0, 0, 0) - U(0, 0, 1);
    const double v1a = V(0, -1, 0) + V(1, -1, 0), v22 = V(0, 0, 0) + V(1, 0, 0);
    const double w1a = W(0, 0, -1) + W(1, 0, -1), w22 = W(0, 0, 0) + W(1, 0, 0);
    donor_adj(-0.25 / A.template hx<-1>() * -g, u1a, u1a, u22, u3, u3, u4, bu1a, bu1a, bu22, bu3, bu3, bu4);
    donor_adj(-0.25 / A.template hy<-1>() * -g, v1a, u5, u6, v22, u7, u8, bv1a, bu5, bu6, bv22, bu7, bu8);
    donor_adj(-0.25 / A.template hz<-1>() * -g, w1a, u9, u10, w22, u11, u12, bw1a, bu9, bu10, bw22, bu11, bu12);
  }
  // v_in, w_in (third order) and the sums / differences
  BV(1, 0, 0) += 0.5 * (1.0 - rx) * bvin; BV(0, 0, 0) += 0.5 * rx * bvin;
  BV(1, -1, 0) += 0.5 * (1.0 - rx) * bvin; BV(0, -1, 0) += 0.5 * rx * bvin;
  BW(1, 0, 0) += 0.5 * (1.0 - rx) * bwin; BW(0, 0, 0) += 0.5 * rx * bwin;
  BW(1, 0, -1) += 0.5 * (1.0 - rx) * bwin; BW(0, 0, -1) += 0.5 * rx * bwin;
  BU(-1, 0, 0) += bu1a + bu22; BU(0, 0, 0) += bu1a - bu22 + bu3 + bu4; BU(1, 0, 0) += bu3 - bu4;
  BU(0, -1, 0) += bu5 + bu6; BU(0, 0, 0) += bu5 - bu6 + bu7 + bu8; BU(0, 1, 0) += bu7 - bu8;
  BU(0, 0, -1) += bu9 + bu10; BU(0, 0, 0) += bu9 - bu10 + bu11 + bu12; BU(0, 0, 1) += bu11 - bu12;
  BV(0, -1, 0) += bv1a; BV(1, -1, 0) += bv1a; BV(0, 0, 0) += bv22; BV(1, 0, 0) += bv22;
  BW(0, 0, -1) += bw1a; BW(1, 0, -1) += bw1a; BW(0, 0, 0) += bw22; BW(1, 0, 0) += bw22;
}

template <bool P2, class Acc>
__device__ inline void mom_v_adj(const MomArgs &a, const Acc &A, const MomCell &m, int i, int j, int k,
                                 bool near_ok, double lam, double bar[49]) {
  const double dxr = m.dxr, dx = m.dx, dxl = m.dxl, dye = m.dye, dy = m.dy, dzt = m.dzt, dz = m.dz, dzb = m.dzb;
  const double dx2xr = dx + dxr, dx2xl = dx + dxl, dy2ye = dy + dye, dz2zt = dz + dzt, dz2zb = dz + dzb;
  const int nx = a.d.nx, ny = a.d.ny, nz = a.d.nz;
  const double g = a.dt * lam;
  BV(0, 0, 0) += lam;
  const double qp = g / (0.5 * (dye + dy));
  BP(0, 0, 0) += qp;
  BP(0, 1, 0) -= qp;
  const double bv = g / a.re;
  const double ax = gdiv<P2>(2.0, 0.5 * (dx2xr + dx2xl)), az = gdiv<P2>(2.0, 0.5 * (dz2zt + dz2zb));
  double bv6 = -ax / (0.5 * dx2xr) * bv, bv4 = ax / (0.5 * dx2xl) * bv;
  double bv10 = -gdiv<P2>(2.0, dy2ye) / dye * bv, bv8 = gdiv<P2>(2.0, dy2ye) / dy * bv;
  double bv14 = -az / (0.5 * dz2zt) * bv, bv12 = az / (0.5 * dz2zb) * bv;
  double bu13 = 0, bu14 = 0, bv3 = 0, bv5 = 0, bv7 = 0, bv9 = 0, bv11 = 0, bv13 = 0, bw3 = 0, bw4 = 0;
  double buin = 0, bwin = 0;
  const double ry = gdiv<P2>(dye, dy2ye);
  if (i > 2 && i < nx + 1 && j > 2 && j < ny && k > 2 && k < nz + 1 && near_ok) {
    const double ddx = 0.5 * (A.template hx<-1>() + A.template hx<-2>()), ddxr = 0.5 * (A.template hx<-1>() + A.template hx<0>());
    const double ddz = 0.5 * (A.template hz<-1>() + A.template hz<-2>()), ddzr = 0.5 * (A.template hz<-1>() + A.template hz<0>());
    const double uv_e = U(0, 1, 0) + ry * (U(0, 0, 0) - U(0, 1, 0));
    const double uv_w = U(-1, 1, 0) + ry * (U(-1, 0, 0) - U(-1, 1, 0));
    const double u_in = 0.5 * (uv_e + uv_w);
    const double wv_n = W(0, 1, 0) + ry * (W(0, 0, 0) - W(0, 1, 0));
    const double wv_s = W(0, 1, -1) + ry * (W(0, 0, -1) - W(0, 1, -1));
    const double w_in = 0.5 * (wv_n + wv_s);
    double bf[5];
    buin = upwind3r_adj(u_in, A.template xa<BTB_UW>(), A.template xk<BTB_UW>(), V(2, 0, 0), V(1, 0, 0), V(0, 0, 0),
                        V(-1, 0, 0), V(-2, 0, 0), ddxr, ddx, -g, bf);
    BV(2, 0, 0) += bf[0]; BV(1, 0, 0) += bf[1]; BV(0, 0, 0) += bf[2]; BV(-1, 0, 0) += bf[3]; BV(-2, 0, 0) += bf[4];
    BV(0, 0, 0) += upwind3r_adj(V(0, 0, 0), A.template ya<BTB_VV>(), A.template yk<BTB_VV>(), V(0, 2, 0), V(0, 1, 0),
                                V(0, 0, 0), V(0, -1, 0), V(0, -2, 0), A.template hy<0>(), A.template hy<-1>(), -g, bf);
    BV(0, 2, 0) += bf[0]; BV(0, 1, 0) += bf[1]; BV(0, 0, 0) += bf[2]; BV(0, -1, 0) += bf[3]; BV(0, -2, 0) += bf[4];
    bwin = upwind3r_adj(w_in, A.template za<BTB_WU>(), A.template zk<BTB_WU>(), V(0, 0, 2), V(0, 0, 1), V(0, 0, 0),
                        V(0, 0, -1), V(0, 0, -2), ddzr, ddz, -g, bf);
    BV(0, 0, 2) += bf[0]; BV(0, 0, 1) += bf[1]; BV(0, 0, 0) += bf[2]; BV(0, 0, -1) += bf[3]; BV(0, 0, -2) += bf[4];
  } else {
    const double u13 = U(-1, 0, 0) + U(-1, 1, 0), u14 = U(0, 0, 0) + U(0, 1, 0);
    const double v3 = V(-1, 0, 0) + V(0, 0, 0), v4 = V(-1, 0, 0) - V(0, 0, 0);
    const double v5 = V(0, 0, 0) + V(1, 0, 0), v6 = V(0, 0, 0) - V(1, 0, 0);
    const double v7 = V(0, -1, 0) + V(0, 0, 0), v8 = V(0, -1, 0) - V(0, 0, 0);
    const double v9 = V(0, 0, 0) + V(0, 1, 0), v10 = V(0, 0, 0) - V(0, 1, 0);
    const double v11 = V(0, 0, -1) + V(0, 0, 0), v12 = V(0, 0, -1) - V(0, 0, 0);
    const double v13 = V(0, 0, 0) + V(0, 0, 1), v14 = V(0, 0, 0) - V(0, 0, 1);
    const double w3 = W(0, 0, -1) + W(0, 1, -1), w4 = W(0, 0, 0) + W(0, 1, 0);
    donor_adj(-0.25 / A.template hx<-1>() * -g, u13, v3, v4, u14, v5, v6, bu13, bv3, bv4, bu14, bv5, bv6);
    donor_adj(-0.25 / A.template hy<-1>() * -g, v7, v7, v8, v9, v9, v10, bv7, bv7, bv8, bv9, bv9, bv10);
    donor_adj(-0.25 / A.template hz<-1>() * -g, w3, v11, v12, w4, v13, v14, bw3, bv11, bv12, bw4, bv13, bv14);
  }
  BU(0, 1, 0) += 0.5 * (1.0 - ry) * buin; BU(0, 0, 0) += 0.5 * ry * buin;
  BU(-1, 1, 0) += 0.5 * (1.0 - ry) * buin; BU(-1, 0, 0) += 0.5 * ry * buin;
  BW(0, 1, 0) += 0.5 * (1.0 - ry) * bwin; BW(0, 0, 0) += 0.5 * ry * bwin;
  BW(0, 1, -1) += 0.5 * (1.0 - ry) * bwin; BW(0, 0, -1) += 0.5 * ry * bwin;
  BV(-1, 0, 0) += bv3 + bv4; BV(0, 0, 0) += bv3 - bv4 + bv5 + bv6; BV(1, 0, 0) += bv5 - bv6;
  BV(0, -1, 0) += bv7 + bv8; BV(0, 0, 0) += bv7 - bv8 + bv9 + bv10; BV(0, 1, 0) += bv9 - bv10;
  BV(0, 0, -1) += bv11 + bv12; BV(0, 0, 0) += bv11 - bv12 + bv13 + bv14; BV(0, 0, 1) += bv13 - bv14;
  BU(-1, 0, 0) += bu13; BU(-1, 1, 0) += bu13; BU(0, 0, 0) += bu14; BU(0, 1, 0) += bu14;
  BW(0, 0, -1) += bw3; BW(0, 1, -1) += bw3; BW(0, 0, 0) += bw4; BW(0, 1, 0) += bw4;
}

template <bool P2, class Acc>
__device__ inline void mom_w_adj(const MomArgs &a, const Acc &A, const MomCell &m, int i, int j, int k,
                                 bool near_ok, double lam, double bar[49]) {
  const double dxr = m.dxr, dx = m.dx, dxl = m.dxl, dye = m.dye, dy = m.dy, dyw = m.dyw, dzt = m.dzt, dz = m.dz;
  const double dx2xr = dx + dxr, dx2xl = dx + dxl, dy2ye = dy + dye, dy2yw = dy + dyw, dz2zt = dz + dzt;
  const int nx = a.d.nx, ny = a.d.ny, nz = a.d.nz;
  const double g = a.dt * lam;
  BW(0, 0, 0) += lam;
  const double qp = g / (0.5 * (dzt + dz));
  BP(0, 0, 0) += qp;
  BP(0, 0, 1) -= qp;
  const double bv = g / a.re;
  const double ax = gdiv<P2>(2.0, 0.5 * (dx2xr + dx2xl)), ay = gdiv<P2>(2.0, 0.5 * (dy2ye + dy2yw));
  double bw8 = -ax / (0.5 * dx2xr) * bv, bw6 = ax / (0.5 * dx2xl) * bv;
  double bw12 = -ay / (0.5 * dy2ye) * bv, bw10 = ay / (0.5 * dy2yw) * bv;
  double bw16 = -gdiv<P2>(2.0, dz2zt) / dzt * bv, bw14 = gdiv<P2>(2.0, dz2zt) / dz * bv;
  double bu15 = 0, bu16 = 0, bv15 = 0, bv16 = 0, bw5 = 0, bw7 = 0, bw9 = 0, bw11 = 0, bw13 = 0, bw15 = 0;
  double buin = 0, bvin = 0;
  const double rz = gdiv<P2>(dzt, dz2zt);
  if (i > 2 && i < nx + 1 && j > 2 && j < ny + 1 && k > 2 && k < nz && near_ok) {
    const double ddx = 0.5 * (A.template hx<-1>() + A.template hx<-2>()), ddxr = 0.5 * (A.template hx<-1>() + A.template hx<0>());
    const double ddy = 0.5 * (A.template hy<-1>() + A.template hy<-2>()), ddye = 0.5 * (A.template hy<-1>() + A.template hy<0>());
    const double uw_e = U(0, 0, 1) + rz * (U(0, 0, 0) - U(0, 0, 1));
    const double uw_w = U(-1, 0, 1) + rz * (U(-1, 0, 0) - U(-1, 0, 1));
    const double u_in = 0.5 * (uw_e + uw_w);
    const double vw_n = V(0, 0, 1) + rz * (V(0, 0, 0) - V(0, 0, 1));
    const double vw_s = V(0, -1, 1) + rz * (V(0, -1, 0) - V(0, -1, 1));
    const double v_in = 0.5 * (vw_n + vw_s);
    double bf[5];
    buin = upwind3r_adj(u_in, A.template xa<BTB_UW>(), A.template xk<BTB_UW>(), W(2, 0, 0), W(1, 0, 0), W(0, 0, 0),
                        W(-1, 0, 0), W(-2, 0, 0), ddxr, ddx, -g, bf);
    BW(2, 0, 0) += bf[0]; BW(1, 0, 0) += bf[1]; BW(0, 0, 0) += bf[2]; BW(-1, 0, 0) += bf[3]; BW(-2, 0, 0) += bf[4];
    bvin = upwind3r_adj(v_in, A.template ya<BTB_VW>(), A.template yk<BTB_VW>(), W(0, 2, 0), W(0, 1, 0), W(0, 0, 0),
                        W(0, -1, 0), W(0, -2, 0), ddye, ddy, -g, bf);
    BW(0, 2, 0) += bf[0]; BW(0, 1, 0) += bf[1]; BW(0, 0, 0) += bf[2]; BW(0, -1, 0) += bf[3]; BW(0, -2, 0) += bf[4];
    BW(0, 0, 0) += upwind3r_adj(W(0, 0, 0), A.template za<BTB_WW>(), A.template zk<BTB_WW>(), W(0, 0, 2), W(0, 0, 1),
                                W(0, 0, 0), W(0, 0, -1), W(0, 0, -2), A.template hz<0>(), A.template hz<-1>(), -g, bf);
    BW(0, 0, 2) += bf[0]; BW(0, 0, 1) += bf[1]; BW(0, 0, 0) += bf[2]; BW(0, 0, -1) += bf[3]; BW(0, 0, -2) += bf[4];
  } else {
    const double u15 = U(-1, 0, 0) + U(-1, 0, 1), u16 = U(0, 0, 0) + U(0, 0, 1);
    const double v15 = V(0, -1, 0) + V(0, -1, 1), v16 = V(0, 0, 0) + V(0, 0, 1);
    const double w5 = W(-1, 0, 0) + W(0, 0, 0), w6 = W(-1, 0, 0) - W(0, 0, 0);
    const double w7 = W(0, 0, 0) + W(1, 0, 0), w8 = W(0, 0, 0) - W(1, 0, 0);
    const double w9 = W(0, -1, 0) + W(0, 0, 0), w10 = W(0, -1, 0) - W(0, 0, 0);
    const double w11 = W(0, 0, 0) + W(0, 1, 0), w12 = W(0, 0, 0) - W(0, 1, 0);
    const double w13 = W(0, 0, -1) + W(0, 0, 0), w14 = W(0, 0, -1) - W(0, 0, 0);
    const double w15 = W(0, 0, 0) + W(0, 0, 1), w16 = W(0, 0, 0) - W(0, 0, 1);
    donor_adj(-0.5 / A.template hx<-1>() * -g, u15, w5, w6, u16, w7, w8, bu15, bw5, bw6, bu16, bw7, bw8);
    donor_adj(-0.5 / A.template hy<-1>() * -g, v15, w9, w10, v16, w11, w12, bv15, bw9, bw10, bv16, bw11, bw12);
    donor_adj(-0.5 / A.template hz<-1>() * -g, w13, w13, w14, w15, w15, w16, bw13, bw13, bw14, bw15, bw15, bw16);
  }
  BU(0, 0, 1) += 0.5 * (1.0 - rz) * buin; BU(0, 0, 0) += 0.5 * rz * buin;
  BU(-1, 0, 1) += 0.5 * (1.0 - rz) * buin; BU(-1, 0, 0) += 0.5 * rz * buin;
  BV(0, 0, 1) += 0.5 * (1.0 - rz) * bvin; BV(0, 0, 0) += 0.5 * rz * bvin;
  BV(0, -1, 1) += 0.5 * (1.0 - rz) * bvin; BV(0, -1, 0) += 0.5 * rz * bvin;
  BW(-1, 0, 0) += bw5 + bw6; BW(0, 0, 0) += bw5 - bw6 + bw7 + bw8; BW(1, 0, 0) += bw7 - bw8;
  BW(0, -1, 0) += bw9 + bw10; BW(0, 0, 0) += bw9 - bw10 + bw11 + bw12; BW(0, 1, 0) += bw11 - bw12;
  BW(0, 0, -1) += bw13 + bw14; BW(0, 0, 0) += bw13 - bw14 + bw15 + bw16; BW(0, 0, 1) += bw15 - bw16;
  BU(-1, 0, 0) += bu15; BU(-1, 0, 1) += bu15; BU(0, 0, 0) += bu16; BU(0, 0, 1) += bu16;
  BV(0, -1, 0) += bv15; BV(0, -1, 1) += bv15; BV(0, 0, 0) += bv16; BV(0, 0, 1) += bv16;
}
#undef BU
#undef BV
#undef BW
#undef BP
#undef U
#undef V
#undef W
#undef P

__device__ inline bool mom_fluid(const MomArgs &a, int64_t q, int &i, int &j, int &k, int64_t &c) {
  const Dims &d = a.d;
  if (q >= (int64_t)d.nx * d.ny * d.nz) return false;
  i = 2 + (int)(q % d.nx);
  j = 2 + (int)((q / d.nx) % d.ny);
  k = 2 + (int)(q / ((int64_t)d.nx * d.ny));
  c = ix3(d, i, j, k);
  return a.cell[c] == 0;
}

template <bool P2>
__global__ void __launch_bounds__(256) k_momentum_tlm(MomArgs a, Fields F, Fields T, double *__restrict__ tut,
                                                      double *__restrict__ tvt, double *__restrict__ twt,
                                                      const uint8_t *__restrict__ deferred) {
  int i, j, k;
  int64_t c;
  if (!mom_fluid(a, (int64_t)blockIdx.x * blockDim.x + threadIdx.x, i, j, k, c)) return;
  if (deferred && deferred[c]) return;
  const MomCell m = mom_cell(a, i, j, k);
  const bool near_ok = mom_near_ok(a, i, j, k);
  DAcc A;
  A.fu = F.u; A.fv = F.v; A.fw = F.w; A.fp = F.p;
  A.c = m.c; A.sy = m.sy; A.sz = m.sz; A.a = &a; A.i = i; A.j = j; A.k = k;
  A.tu = T.u; A.tv = T.v; A.tw = T.w; A.tp = T.p;
  tut[c] = mom_u<P2>(a, A, m, i, j, k, near_ok).t;
  tvt[c] = mom_v<P2>(a, A, m, i, j, k, near_ok).t;
  twt[c] = mom_w<P2>(a, A, m, i, j, k, near_ok).t;
}

// lambda of ut, vt, wt at a fluid cell is consumed (the forward overwrote
// them) and the reverse of the component functions scatters it, atomically,
// into the adjoints of the operand positions
template <bool P2>
__global__ void __launch_bounds__(256) k_momentum_adj(MomArgs a, Fields F, double *__restrict__ lut,
                                                      double *__restrict__ lvt, double *__restrict__ lwt,
                                                      double *__restrict__ lu, double *__restrict__ lv,
                                                      double *__restrict__ lw, double *__restrict__ lp,
                                                      const uint8_t *__restrict__ deferred) {
  int i, j, k;
  int64_t c;
  if (!mom_fluid(a, (int64_t)blockIdx.x * blockDim.x + threadIdx.x, i, j, k, c)) return;
  if (deferred && deferred[c]) return;
  const double gu = lut[c], gv = lvt[c], gw = lwt[c];
  lut[c] = 0.0;
  lvt[c] = 0.0;
  lwt[c] = 0.0;
  if (gu == 0.0 && gv == 0.0 && gw == 0.0) return;
  const MomCell m = mom_cell(a, i, j, k);
  const bool near_ok = mom_near_ok(a, i, j, k);
  const GAcc A{F.u, F.v, F.w, F.p, m.c, m.sy, m.sz, &a, i, j, k};
  double bar[49];
#pragma unroll
  for (int q = 0; q < 49; ++q) bar[q] = 0.0;
  mom_u_adj<P2>(a, A, m, i, j, k, near_ok, gu, bar);
  mom_v_adj<P2>(a, A, m, i, j, k, near_ok, gv, bar);
  mom_w_adj<P2>(a, A, m, i, j, k, near_ok, gw, bar);
  double *lam[4] = {lu, lv, lw, lp};
#pragma unroll
  for (int q = 0; q < 49; ++q) {
    const Seed s = kSeeds[q];
    if (bar[q] != 0.0) atomicAdd(&lam[s.f][c + s.i + (int64_t)s.j * m.sy + (int64_t)s.k * m.sz], bar[q]);
  }
}

__global__ void k_fill_u8(uint8_t *__restrict__ x, int64_t n) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n) x[q] = 0;
}

__global__ void k_fill_f64(double *__restrict__ x, int64_t n, double v) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n) x[q] = v;
}

// The thin-surface pass's deferred cells (k_ts_parallel): every operand is
// read from where the ordered pass lets the cell read it, a field entry or a
// ghost entry (TsAcc::src), and its tangent or adjoint goes to the same place.
// Ghost tangents / adjoints: gh (BGH_N arrays indexed like the ghosts).
struct TsSens {
  double *f[4];
  double *gh[BGH_N];
};

// TsAcc with the tangents of the operands beside them (same table layout)
struct TsDAcc {
  TsAcc base;
  const double *tval;
  template <int F, int di, int dj, int dk>
  __device__ Dual at() const {
    const int64_t e = (int64_t)sidx<F, di, dj, dk>() * base.n + base.q;
    return {base.val[e], tval[e]};
  }
  template <int di, int dj, int dk> __device__ Dual u() const { return at<0, di, dj, dk>(); }
  template <int di, int dj, int dk> __device__ Dual v() const { return at<1, di, dj, dk>(); }
  template <int di, int dj, int dk> __device__ Dual w() const { return at<2, di, dj, dk>(); }
  template <int di, int dj, int dk> __device__ Dual p() const { return at<3, di, dj, dk>(); }
  template <int C> __device__ Dual own() const { return at<C, 0, 0, 0>(); }
  template <int o> __device__ double hx() const { return base.template hx<o>(); }
  template <int o> __device__ double hy() const { return base.template hy<o>(); }
  template <int o> __device__ double hz() const { return base.template hz<o>(); }
  template <int T> __device__ const double *xa() const { return base.template xa<T>(); }
  template <int T> __device__ const double *xk() const { return base.template xk<T>(); }
  template <int T> __device__ const double *ya() const { return base.template ya<T>(); }
  template <int T> __device__ const double *yk() const { return base.template yk<T>(); }
  template <int T> __device__ const double *za() const { return base.template za<T>(); }
  template <int T> __device__ const double *zk() const { return base.template zk<T>(); }
  // no reciprocal rows: the stencil divides
  template <int T> __device__ std::nullptr_t xr() const { return nullptr; }
  template <int T> __device__ std::nullptr_t yr() const { return nullptr; }
  template <int T> __device__ std::nullptr_t zr() const { return nullptr; }
  __device__ int *bad() const { return nullptr; }
};

__device__ inline void ts_cell(const Dims &d, const int *list, int q, int &i, int &j, int &k) {
  const int lin = list[q];
  i = 2 + lin % d.nx;
  j = 2 + (lin / d.nx) % d.ny;
  k = 2 + lin / (d.nx * d.ny);
}

__global__ void __launch_bounds__(256) k_ts_tlm(MomArgs a, const double *__restrict__ val,
                                                const double *__restrict__ tval, const int *__restrict__ list,
                                                int n_list, double *__restrict__ tut, double *__restrict__ tvt,
                                                double *__restrict__ twt) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n_list) return;
  int i, j, k;
  ts_cell(a.d, list, q, i, j, k);
  const MomCell m = mom_cell(a, i, j, k);
  const bool near_ok = mom_near_ok(a, i, j, k);
  const TsDAcc A{{val, n_list, &a, q, i, j, k}, tval};
  tut[m.c] = mom_u<false>(a, A, m, i, j, k, near_ok).t;
  tvt[m.c] = mom_v<false>(a, A, m, i, j, k, near_ok).t;
  twt[m.c] = mom_w<false>(a, A, m, i, j, k, near_ok).t;
}

// the operand adjoints of each deferred cell into bars (the operand table's
// layout); k_ts_scatter adds them to their sources
__global__ void __launch_bounds__(256) k_ts_adj(MomArgs a, const double *__restrict__ val, double *__restrict__ bars,
                                                const int *__restrict__ list, int n_list, double *__restrict__ lut,
                                                double *__restrict__ lvt, double *__restrict__ lwt) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n_list) return;
  int i, j, k;
  ts_cell(a.d, list, q, i, j, k);
  const MomCell m = mom_cell(a, i, j, k);
  const double gu = lut[m.c], gv = lvt[m.c], gw = lwt[m.c];
  lut[m.c] = 0.0;
  lvt[m.c] = 0.0;
  lwt[m.c] = 0.0;
  double bar[49];
#pragma unroll
  for (int s = 0; s < 49; ++s) bar[s] = 0.0;
  if (gu != 0.0 || gv != 0.0 || gw != 0.0) {
    const bool near_ok = mom_near_ok(a, i, j, k);
    const TsAcc A{val, n_list, &a, q, i, j, k};
    mom_u_adj<false>(a, A, m, i, j, k, near_ok, gu, bar);
    mom_v_adj<false>(a, A, m, i, j, k, near_ok, gv, bar);
    mom_w_adj<false>(a, A, m, i, j, k, near_ok, gw, bar);
  }
#pragma unroll
  for (int s = 0; s < 49; ++s) bars[(int64_t)s * n_list + q] = bar[s];
}

__global__ void k_ts_scatter(const int64_t *__restrict__ src, const double *__restrict__ bars, int64_t n,
                             TsArrays<double> L) {
  const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= n) return;
  const double b = bars[gid];
  if (b == 0.0) return;
  const int64_t w = src[gid];
  atomicAdd(L.of((int)(w & 31)) + (w >> 5), b);
}

// transpose of k_ts_restore: the adjoint at every restored position goes to
// the ghost entry its last writer stored there
__global__ void k_ts_restore_adj(Dims d, TsSens L, const int *__restrict__ list, int n_list, TsView tv) {
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  const int q = gid / 36, slot = gid % 36;
  if (q >= n_list) return;
  const int dq = slot / 6;
  if (!((tv.adjm[q] >> dq) & 1u)) return;
  int i, j, k;
  ts_cell(d, list, q, i, j, k);
  const Target T = kTargets[dq][slot % 6];
  const int pi = i + T.di, pj = j + T.dj, pk = k + T.dk;
  if (pi < 1 || pj < 1 || pk < 1 || pi > d.NX || pj > d.NY || pk > d.NZ) return;
  bool last = true;
  ts_writers(d, tv, T.f, pi, pj, pk, [&](int qw, int s2) {
    if (qw > q || (qw == q && s2 > slot)) last = false;
  });
  if (!last) return;
  double *lf = &L.f[T.f][ix3(d, pi, pj, pk)];
  const double g = *lf;
  *lf = 0.0;
  if (g != 0.0) atomicAdd(&L.gh[T.gt][tv.n_ts - 1], g);
}

}  // namespace

int compact_mask_lin(Block &b, const uint8_t *mask, int **out, int64_t *cap, int64_t *count);

// The deferred cells of the thin-surface pass and the structures of its
// parallel form, as biocfd_ns_momentum builds them (geometry only); cnt 0
// and deferred null without TS cells.  ts_release clears the cell map.
static int ts_prepare(Block &b, const MomArgs &a, TsView &tv, const int *&list, int64_t &cnt,
                      const uint8_t *&deferred) {
  cnt = 0;
  deferred = nullptr;
  list = nullptr;
  const int n_ts = (int)b.counts[BC_TS];
  if (n_ts == 0) return 0;
  IbData &ib = b.ib;
  hipStream_t s = rt().stream;
  if (grow(&ib.ts_mask, &ib.cap_ts_mask, 3 * b.n2)) return -1;
  uint8_t *touched = ib.ts_mask, *dfr = ib.ts_mask + b.n2, *near = ib.ts_mask + 2 * b.n2;
  const dim3 blk(64, 4, 1), grd(div_up(b.d.nx, 64), div_up(b.d.ny, 4), (unsigned)b.d.nz);
  hipLaunchKernelGGL(k_fill_u8, dim3(div_up(3 * b.n2, 256)), dim3(256), 0, s, ib.ts_mask, 3 * b.n2);
  hipLaunchKernelGGL(k_ts_touch, grd, blk, 0, s, b.d, b.tag[BT_CELL], b.tag[BT_CELL2], touched);
  hipLaunchKernelGGL(k_ts_near, dim3(div_up(b.d.NX, 64), div_up(b.d.NY, 4), (unsigned)b.d.NZ), blk, 0, s, b.d,
                     touched, near);
  hipLaunchKernelGGL(k_ts_defer, grd, blk, 0, s, b.d, b.tag[BT_CELL], b.tag[BT_CELL2], near, dfr);
  BCFD_CHECK_LAUNCH();
  if (compact_mask_lin(b, dfr, &ib.ts_list, &ib.cap_ts_list, &cnt)) return -1;
  deferred = dfr;
  list = ib.ts_list;
  if (cnt == 0) return 0;
  TsArgs ts{};
  ts.ts_ijk = ib.ts_ijk; ts.cap_ts = ib.cap_ts; ts.index_ts = ib.index_ts;
  for (int q = 0; q < BN_NNEL; ++q) ts.nel[q] = ib.nel[q];
  for (int q = 0; q < BE_NELEM; ++q) ts.elem[q] = ib.elem[q];
  for (int q = 0; q < BGH_N; ++q) ts.ghost[q] = ib.ghost[q];
  ts.x1 = b.g[BG_X1]; ts.y1 = b.g[BG_Y1]; ts.z1 = b.g[BG_Z1];
  ts.xp = b.g[BG_XP]; ts.yp = b.g[BG_YP]; ts.zp = b.g[BG_ZP];
  ts.n_ts = n_ts;
  if (grow(&ib.ts_best, &ib.cap_ts_best, 36 * cnt)) return -1;
  hipLaunchKernelGGL(k_ts_best, dim3((unsigned)cnt), dim3(256), 0, s, a, list, (int)cnt, ts, ib.ts_best);
  const int *old_map = ib.ts_qmap;
  if (grow(&ib.ts_qmap, &ib.cap_ts_qmap, b.n2) || grow(&ib.ts_adjm, &ib.cap_ts_adjm, cnt)) return -1;
  if (ib.ts_qmap != old_map)
    hipLaunchKernelGGL(k_fill_int, dim3(div_up(ib.cap_ts_qmap, 256)), dim3(256), 0, s, ib.ts_qmap, ib.cap_ts_qmap, -1);
  tv = TsView{};
  tv.qmap = ib.ts_qmap;
  tv.adjm = ib.ts_adjm;
  tv.best = ib.ts_best;
  tv.touched = touched;
  for (int g = 0; g < BGH_N; ++g) tv.ghost[g] = ib.ghost[g];
  tv.n_ts = n_ts;
  hipLaunchKernelGGL(k_ts_prep, dim3(div_up(cnt, 256)), dim3(256), 0, s, b.d, b.tag[BT_CELL2], list, (int)cnt,
                     ib.ts_qmap, ib.ts_adjm, 0);
  BCFD_CHECK_LAUNCH();
  return 0;
}

static int ts_release(Block &b, const int *list, int64_t cnt) {
  if (cnt == 0) return 0;
  hipLaunchKernelGGL(k_ts_prep, dim3(div_up(cnt, 256)), dim3(256), 0, rt().stream, b.d, b.tag[BT_CELL2], list,
                     (int)cnt, b.ib.ts_qmap, b.ib.ts_adjm, 1);
  BCFD_CHECK_LAUNCH();
  return 0;
}

// the ghost sensitivities, at least as long as the TS arrays (new entries 0)
int ad_ghosts(Block &b) {
  AdWork &w = b.ad;
  const int64_t need = b.ib.cap_ts > 0 ? b.ib.cap_ts : 1;
  if (w.gh[0] && w.cap_gh >= need) return 0;
  for (int q = 0; q < BGH_N; ++q) {
    if (dev_alloc(&w.gh[q], need)) return -1;
    hipLaunchKernelGGL(k_fill_f64, dim3(div_up(need, 256)), dim3(256), 0, rt().stream, w.gh[q], need, 0.0);
  }
  w.cap_gh = need;
  BCFD_CHECK_LAUNCH();
  return 0;
}

static int momentum_args(Block &b, double dt, double re, MomArgs &a) {
  for (int t = 0; t < BTB_N; ++t)
    if (!b.ca[t]) { set_error("non_uni_coeff has not been called"); return -1; }
  if (!b.g[BG_DX]) { set_error("grids not uploaded"); return -1; }
  a.d = b.d;
  a.cell = b.tag[BT_CELL];
  a.cell2 = b.tag[BT_CELL2];
  a.dx = b.g[BG_DX]; a.dy = b.g[BG_DY]; a.dz = b.g[BG_DZ];
  for (int t = 0; t < BTB_N; ++t) {
    a.ca[t] = b.ca[t]; a.ck[t] = b.ck[t];
    a.ca32[t] = b.ca32[t]; a.ck32[t] = b.ck32[t];
  }
  a.dx32 = b.h32[0]; a.dy32 = b.h32[1]; a.dz32 = b.h32[2];
  for (int t = 0; t < BTB_N; ++t) a.rr[t] = b.rr[t];
  a.dt = dt;
  a.re = re;
  a.rre = 1.0 / re;
  return 0;
}

int momentum_tlm(Block &b, double dt, double re) {
  MomArgs a;
  if (momentum_args(b, dt, re, a)) return -1;
  TsView tv;
  const int *list = nullptr;
  const uint8_t *deferred = nullptr;
  int64_t cnt = 0;
  if (ts_prepare(b, a, tv, list, cnt, deferred) || ad_ghosts(b)) return -1;
  hipStream_t s = rt().stream;
  double **t = b.ad.s;
  const Fields F{b.f[BF_U], b.f[BF_V], b.f[BF_W], b.f[BF_P]};
  const Fields T{t[BF_U], t[BF_V], t[BF_W], t[BF_P]};
  const unsigned g = div_up((int64_t)b.d.nx * b.d.ny * b.d.nz, 256);
  if (g_mom_pow2 && pow2_grid(b))
    hipLaunchKernelGGL(k_momentum_tlm<true>, dim3(g), dim3(256), 0, s, a, F, T, t[BF_UT], t[BF_VT], t[BF_WT], deferred);
  else
    hipLaunchKernelGGL(k_momentum_tlm<false>, dim3(g), dim3(256), 0, s, a, F, T, t[BF_UT], t[BF_VT], t[BF_WT],
                       deferred);
  if (cnt > 0) {
    TsArrays<const double> fa{{b.f[BF_U], b.f[BF_V], b.f[BF_W], b.f[BF_P]}};
    TsArrays<const double> ta{{t[BF_U], t[BF_V], t[BF_W], t[BF_P]}};
    for (int g = 0; g < BGH_N; ++g) {
      fa.a[4 + g] = b.ib.ghost[g];
      ta.a[4 + g] = b.ad.gh[g];
    }
    double *val = nullptr;
    if (ts_operands(b, list, cnt, tv, fa, &val)) return -1;
    double *tval = val + 49 * cnt;
    hipLaunchKernelGGL(k_ts_gather, dim3(div_up(49 * cnt, 256)), dim3(256), 0, s, b.ib.ts_src, 49 * cnt, ta, tval);
    hipLaunchKernelGGL(k_ts_tlm, dim3(div_up(cnt, 256)), dim3(256), 0, s, a, val, tval, list, (int)cnt, t[BF_UT],
                       t[BF_VT], t[BF_WT]);
    // the pass leaves each restored position holding its last writer's
    // restore ghost: so do the tangents
    TsView tt = tv;
    for (int q = 0; q < BGH_N; ++q) tt.ghost[q] = b.ad.gh[q];
    hipLaunchKernelGGL(k_ts_restore, dim3(div_up(36 * cnt, 256)), dim3(256), 0, s, b.d, t[BF_U], t[BF_V], t[BF_W],
                       t[BF_P], list, (int)cnt, tt);
  }
  BCFD_CHECK_LAUNCH();
  return ts_release(b, list, cnt);
}

int momentum_adj(Block &b, double dt, double re) {
  MomArgs a;
  if (momentum_args(b, dt, re, a)) return -1;
  TsView tv;
  const int *list = nullptr;
  const uint8_t *deferred = nullptr;
  int64_t cnt = 0;
  if (ts_prepare(b, a, tv, list, cnt, deferred) || ad_ghosts(b)) return -1;
  hipStream_t s = rt().stream;
  double **l = b.ad.s;
  const Fields F{b.f[BF_U], b.f[BF_V], b.f[BF_W], b.f[BF_P]};
  TsSens ls{{l[BF_U], l[BF_V], l[BF_W], l[BF_P]}, {}};
  for (int q = 0; q < BGH_N; ++q) ls.gh[q] = b.ad.gh[q];
  // reverse order of the forward: the restores, the deferred cells, the dense pass
  if (cnt > 0) {
    hipLaunchKernelGGL(k_ts_restore_adj, dim3(div_up(36 * cnt, 256)), dim3(256), 0, s, b.d, ls, list, (int)cnt, tv);
    TsArrays<const double> fa{{b.f[BF_U], b.f[BF_V], b.f[BF_W], b.f[BF_P]}};
    TsArrays<double> la{{l[BF_U], l[BF_V], l[BF_W], l[BF_P]}};
    for (int g = 0; g < BGH_N; ++g) {
      fa.a[4 + g] = b.ib.ghost[g];
      la.a[4 + g] = b.ad.gh[g];
    }
    double *val = nullptr;
    if (ts_operands(b, list, cnt, tv, fa, &val)) return -1;
    double *bars = val + 49 * cnt;
    hipLaunchKernelGGL(k_ts_adj, dim3(div_up(cnt, 256)), dim3(256), 0, s, a, val, bars, list, (int)cnt, l[BF_UT],
                       l[BF_VT], l[BF_WT]);
    hipLaunchKernelGGL(k_ts_scatter, dim3(div_up(49 * cnt, 256)), dim3(256), 0, s, b.ib.ts_src, bars, 49 * cnt, la);
  }
  const unsigned g = div_up((int64_t)b.d.nx * b.d.ny * b.d.nz, 256);
  if (g_mom_pow2 && pow2_grid(b))
    hipLaunchKernelGGL(k_momentum_adj<true>, dim3(g), dim3(256), 0, s, a, F, l[BF_UT], l[BF_VT], l[BF_WT], l[BF_U],
                       l[BF_V], l[BF_W], l[BF_P], deferred);
  else
    hipLaunchKernelGGL(k_momentum_adj<false>, dim3(g), dim3(256), 0, s, a, F, l[BF_UT], l[BF_VT], l[BF_WT], l[BF_U],
                       l[BF_V], l[BF_W], l[BF_P], deferred);
  BCFD_CHECK_LAUNCH();
  return ts_release(b, list, cnt);
}

}  // namespace bcfd
