// tff_kernels.h -- HIP kernels of the force-free step for gfx950.
//
// One thread per cell of the allocated box (interior + NGHOST ghosts),
// x fastest so that a wave64 reads 64 consecutive elements of a row.  Each
// kernel writes every cell of its output: the interior gets the new value,
// ghost cells get the input's ghost value, so an output buffer carries the
// same ghost contents the reference's in-place arrays keep.
#pragma once
#include "tff_cell.h"

namespace tff {

constexpr int BX = 64;  // threads along x: one wave64 per row segment
constexpr int BY = 4;

__device__ __forceinline__ bool interior(const Geom& g, int i, int j, int k) {
  return i >= 0 && i < g.nx && j >= 0 && j < g.ny && k >= 0 && k < g.nz;
}

// ---------------------------------------------------------------------------
// rk3Step (field_solver.F90:35-42): computeRho (:44-63), rk3Push (:65-166) and
// rk3Update (:168-271) fused per cell.  Reads X = (xe, xb) with a one-cell
// stencil, U0 = (ue, ub) at the cell, the precomputed B0 curl term t0 and
// writes Y = c1*U0 + c2*X + c3*L(X) on the interior.  With P == T == S the
// operation order is the reference's, so FP64 and FP32 results are those of
// field_solver.F90 bit for bit (built with -ffp-contract=off).
template <class PR>
__global__ void __launch_bounds__(BX* BY)
    k_rk3(Geom g, V3<const typename PR::S> xe, V3<const typename PR::S> xb,
          V3<const typename PR::S> ue, V3<const typename PR::S> ub,
          V3<const typename PR::S> t0, int has_t0, V3<typename PR::S> ye,
          V3<typename PR::S> yb, typename PR::P cc, typename PR::P c1, typename PR::P c2,
          typename PR::P c3, Absorb<typename PR::P> ab) {
  using S = typename PR::S;
  using T = typename PR::T;
  using P = typename PR::P;
  const int i = blockIdx.x * BX + threadIdx.x - NG;
  const int j = blockIdx.y * BY + threadIdx.y - NG;
  const int k = blockIdx.z - NG;
  if (i > g.nx - 1 + NG || j > g.ny - 1 + NG) return;
  const int64_t c = g.at(i, j, k);
  if (!interior(g, i, j, k)) {
    ye.x[c] = xe.x[c];
    ye.y[c] = xe.y[c];
    ye.z[c] = xe.z[c];
    yb.x[c] = xb.x[c];
    yb.y[c] = xb.y[c];
    yb.z[c] = xb.z[c];
    return;
  }
  const int64_t sy = g.px, sz = g.plane;
  auto off = [=](int a, int b, int d) { return c + a + b * sy + d * sz; };
  auto EX = [&](int a, int b, int d) { return (T)xe.x[off(a, b, d)]; };
  auto EY = [&](int a, int b, int d) { return (T)xe.y[off(a, b, d)]; };
  auto EZ = [&](int a, int b, int d) { return (T)xe.z[off(a, b, d)]; };
  auto BXF = [&](int a, int b, int d) { return (T)xb.x[off(a, b, d)]; };
  auto BYF = [&](int a, int b, int d) { return (T)xb.y[off(a, b, d)]; };
  auto BZF = [&](int a, int b, int d) { return (T)xb.z[off(a, b, d)]; };
  auto RHO = [&](int a, int b, int d) { return rho_node<T>(EX, EY, EZ, a, b, d); };
  const T tx = has_t0 ? (T)t0.x[c] : (T)0;
  const T ty = has_t0 ? (T)t0.y[c] : (T)0;
  const T tz = has_t0 ? (T)t0.z[c] : (T)0;
  P d[6];
  rhs_cell<PR>(EX, EY, EZ, BXF, BYF, BZF, RHO, cc, tx, ty, tz, d);
  P u[6] = {(P)ue.x[c], (P)ue.y[c], (P)ue.z[c], (P)ub.x[c], (P)ub.y[c], (P)ub.z[c]};
  P v[6] = {(P)xe.x[c], (P)xe.y[c], (P)xe.z[c], (P)xb.x[c], (P)xb.y[c], (P)xb.z[c]};
  update_cell<PR>(v, u, d, c1, c2, c3, ab, i, j, k);
  ye.x[c] = (S)v[0];
  ye.y[c] = (S)v[1];
  ye.z[c] = (S)v[2];
  yb.x[c] = (S)v[3];
  yb.y[c] = (S)v[4];
  yb.z[c] = (S)v[5];
}

// ---------------------------------------------------------------------------
// cleanEpar (field_solver.F90:273-335): E <- E - (E.B) B / |B|^2 with the
// staggered interpolants, no TINY (an exact magnetic null gives NaN exactly
// as the reference does).  With `guard` (the fixed mode, whose absorber
// zeroes the fields beyond r_max) the projection is skipped where |B|^2 is
// exactly zero.  Interior only; ghosts carry the input value.
template <class PR>
__global__ void __launch_bounds__(BX* BY)
    k_clean(Geom g, V3<const typename PR::S> xe, V3<const typename PR::S> xb,
            V3<typename PR::S> ye, int guard) {
  using S = typename PR::S;
  using T = typename PR::T;
  using P = typename PR::P;
  const int i = blockIdx.x * BX + threadIdx.x - NG;
  const int j = blockIdx.y * BY + threadIdx.y - NG;
  const int k = blockIdx.z - NG;
  if (i > g.nx - 1 + NG || j > g.ny - 1 + NG) return;
  const int64_t c = g.at(i, j, k);
  if (!interior(g, i, j, k)) {
    ye.x[c] = xe.x[c];
    ye.y[c] = xe.y[c];
    ye.z[c] = xe.z[c];
    return;
  }
  const int64_t sy = g.px, sz = g.plane;
  auto off = [=](int a, int b, int d) { return c + a + b * sy + d * sz; };
  auto EX = [&](int a, int b, int d) { return (T)xe.x[off(a, b, d)]; };
  auto EY = [&](int a, int b, int d) { return (T)xe.y[off(a, b, d)]; };
  auto EZ = [&](int a, int b, int d) { return (T)xe.z[off(a, b, d)]; };
  auto BXF = [&](int a, int b, int d) { return (T)xb.x[off(a, b, d)]; };
  auto BYF = [&](int a, int b, int d) { return (T)xb.y[off(a, b, d)]; };
  auto BZF = [&](int a, int b, int d) { return (T)xb.z[off(a, b, d)]; };
  P ox, oy, oz;
  clean_cell<PR>(EX, EY, EZ, BXF, BYF, BZF, guard, ox, oy, oz);
  ye.x[c] = (S)ox;
  ye.y[c] = (S)oy;
  ye.z[c] = (S)oz;
}

// ---------------------------------------------------------------------------
// The compensated path (TFF_FP32_COMP): each field buffer holds its lo
// words `lo` elements after the hi words (hi x y z, then lo x y z).
//
// rk3Step on pairs: the push evaluates the FP32-bulk right-hand side from
// the hi words (with its FP64 redo near magnetic nulls); the update and the
// absorbing layer carry lo (comp_update, comp_scale).  c1 = 1 - c2, as for
// every SSP-RK3 stage.
template <class PR>
__global__ void __launch_bounds__(BX* BY)
    k_rk3c(Geom g, V3<const float> xe, V3<const float> xb, V3<const float> ue, V3<const float> ub,
           V3<const float> t0, int has_t0, V3<float> ye, V3<float> yb, float cc, float c2, float c3,
           Absorb<float> ab, int64_t lo) {
  const int i = blockIdx.x * BX + threadIdx.x - NG;
  const int j = blockIdx.y * BY + threadIdx.y - NG;
  const int k = blockIdx.z - NG;
  if (i > g.nx - 1 + NG || j > g.ny - 1 + NG) return;
  const int64_t c = g.at(i, j, k);
  const float* xs[6] = {xe.x, xe.y, xe.z, xb.x, xb.y, xb.z};
  float* ys[6] = {ye.x, ye.y, ye.z, yb.x, yb.y, yb.z};
  if (!interior(g, i, j, k)) {
#pragma unroll
    for (int m = 0; m < 6; ++m) {
      ys[m][c] = xs[m][c];
      ys[m][c + lo] = xs[m][c + lo];
    }
    return;
  }
  const int64_t sy = g.px, sz = g.plane;
  auto off = [=](int a, int b, int d) { return c + a + b * sy + d * sz; };
  auto EX = [&](int a, int b, int d) { return xe.x[off(a, b, d)]; };
  auto EY = [&](int a, int b, int d) { return xe.y[off(a, b, d)]; };
  auto EZ = [&](int a, int b, int d) { return xe.z[off(a, b, d)]; };
  auto BXF = [&](int a, int b, int d) { return xb.x[off(a, b, d)]; };
  auto BYF = [&](int a, int b, int d) { return xb.y[off(a, b, d)]; };
  auto BZF = [&](int a, int b, int d) { return xb.z[off(a, b, d)]; };
  auto RHO = [&](int a, int b, int d) { return rho_node<float>(EX, EY, EZ, a, b, d); };
  const float tx = has_t0 ? t0.x[c] : 0.f, ty = has_t0 ? t0.y[c] : 0.f, tz = has_t0 ? t0.z[c] : 0.f;
  typename PR::P d[6];
  rhs_cell<PR>(EX, EY, EZ, BXF, BYF, BZF, RHO, cc, tx, ty, tz, d);
  const float* us[6] = {ue.x, ue.y, ue.z, ub.x, ub.y, ub.z};
#pragma unroll
  for (int m = 0; m < 6; ++m) {
    float yh, yl;
    comp_update(us[m][c], us[m][c + lo], xs[m][c], xs[m][c + lo], (float)d[m], c2, c3, yh, yl);
    if (ab.on) comp_scale(yh, yl, absorb_factor(ab, i, j, k, m));
    ys[m][c] = yh;
    ys[m][c + lo] = yl;
  }
}

// cleanEpar on pairs: the projection in FP64 from hi + lo of every E and B
// word it reads (so the parallel part of lo is removed too), the result
// split back into a pair.  `guard` as in k_clean.
__global__ void __launch_bounds__(BX* BY)
    k_cleanc(Geom g, V3<const float> xe, V3<const float> xb, V3<float> ye, int guard, int64_t lo) {
  const int i = blockIdx.x * BX + threadIdx.x - NG;
  const int j = blockIdx.y * BY + threadIdx.y - NG;
  const int k = blockIdx.z - NG;
  if (i > g.nx - 1 + NG || j > g.ny - 1 + NG) return;
  const int64_t c = g.at(i, j, k);
  const float* xs[3] = {xe.x, xe.y, xe.z};
  float* ys[3] = {ye.x, ye.y, ye.z};
  if (!interior(g, i, j, k)) {
#pragma unroll
    for (int m = 0; m < 3; ++m) {
      ys[m][c] = xs[m][c];
      ys[m][c + lo] = xs[m][c + lo];
    }
    return;
  }
  const int64_t sy = g.px, sz = g.plane;
  auto pair = [=](const float* a, int x, int y, int z) {
    const int64_t o = c + x + y * sy + z * sz;
    return (double)a[o] + (double)a[o + lo];
  };
  auto EX = [&](int a, int b, int d) { return pair(xe.x, a, b, d); };
  auto EY = [&](int a, int b, int d) { return pair(xe.y, a, b, d); };
  auto EZ = [&](int a, int b, int d) { return pair(xe.z, a, b, d); };
  auto BXF = [&](int a, int b, int d) { return pair(xb.x, a, b, d); };
  auto BYF = [&](int a, int b, int d) { return pair(xb.y, a, b, d); };
  auto BZF = [&](int a, int b, int d) { return pair(xb.z, a, b, d); };
  const double o[3] = {clean_point_fp64<0>(EX, EY, EZ, BXF, BYF, BZF, guard),
                       clean_point_fp64<1>(EX, EY, EZ, BXF, BYF, BZF, guard),
                       clean_point_fp64<2>(EX, EY, EZ, BXF, BYF, BZF, guard)};
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    const float h = (float)o[m];
    ys[m][c] = h;
    ys[m][c + lo] = (float)(o[m] - (double)h);
  }
}

// ---------------------------------------------------------------------------
// Ghost shell enumeration: z-slabs (all i, j), then y-slabs (interior k, all
// i), then x-slabs (interior j, k).  Every ghost cell appears exactly once.
// A shell has far fewer than 2^32 cells, so the index arithmetic is 32-bit
// (64-bit division is a long software sequence on the GPU).
__device__ __forceinline__ bool shell_cell(const Geom& g, int64_t t64, int& i, int& j, int& k) {
  const uint32_t sx = (uint32_t)(g.nx + 2 * NG), sy = (uint32_t)(g.ny + 2 * NG);
  const uint32_t nz = (uint32_t)g.nz, ny = (uint32_t)g.ny;
  const uint32_t nzc = 2u * NG * sx * sy;
  const uint32_t nyc = 2u * NG * sx * nz;
  const uint32_t nxc = 2u * NG * ny * nz;
  uint32_t t = (uint32_t)t64;
  if (t < nzc) {
    i = (int)(t % sx) - NG;
    t /= sx;
    j = (int)(t % sy) - NG;
    const int kk = (int)(t / sy);
    k = kk < NG ? kk - NG : g.nz + kk - NG;
    return true;
  }
  t -= nzc;
  if (t < nyc) {
    i = (int)(t % sx) - NG;
    t /= sx;
    k = (int)(t % nz);
    const int jj = (int)(t / nz);
    j = jj < NG ? jj - NG : g.ny + jj - NG;
    return true;
  }
  t -= nyc;
  if (t < nxc) {
    const int ii = (int)(t % (2u * NG));
    t /= (2u * NG);
    j = (int)(t % ny);
    k = (int)(t / ny);
    i = ii < NG ? ii - NG : g.nx + ii - NG;
    return true;
  }
  return false;
}

// exchangeFields on one meshblock (exchange_fields.F90:44-195 with every
// neighbour being the block itself, helpers.F90:22-46): a ghost cell is
// filled iff every axis on which it lies outside the interior is periodic,
// and it takes the value of the wrapped interior cell.
template <class S>
struct FillSet {
  S* a[12];
  int n;
};

template <class S>
__global__ void k_fill(Geom g, int px_, int py_, int pz_, FillSet<S> fs, int64_t nshell) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nshell) return;
  int i, j, k;
  shell_cell(g, t, i, j, k);
  const bool ox = i < 0 || i >= g.nx, oy = j < 0 || j >= g.ny, oz = k < 0 || k >= g.nz;
  if ((ox && !px_) || (oy && !py_) || (oz && !pz_)) return;
  const int wi = i < 0 ? i + g.nx : (i >= g.nx ? i - g.nx : i);
  const int wj = j < 0 ? j + g.ny : (j >= g.ny ? j - g.ny : j);
  const int wk = k < 0 ? k + g.nz : (k >= g.nz ? k - g.nz : k);
  const int64_t d = g.at(i, j, k), s = g.at(wi, wj, wk);
  for (int f = 0; f < fs.n; ++f) fs.a[f][d] = fs.a[f][s];
}

// Curl of the background field, the B0 bracket of rk3Push (:84-89),
// evaluated once: t0 = ((b0y(k)-b0y(k-1)) - (b0z(j)-b0z(j-1))) and cyclic.
template <class S>
__global__ void __launch_bounds__(BX* BY)
    k_b0_curl(Geom g, V3<const S> b0, V3<S> t0, int* nonzero) {
  const int i = blockIdx.x * BX + threadIdx.x;
  const int j = blockIdx.y * BY + threadIdx.y;
  const int k = blockIdx.z;
  if (i >= g.nx || j >= g.ny) return;
  const int64_t c = g.at(i, j, k), sy = g.px, sz = g.plane;
  const S tx = (b0.y[c] - b0.y[c - sz]) - (b0.z[c] - b0.z[c - sy]);
  const S ty = (b0.z[c] - b0.z[c - 1]) - (b0.x[c] - b0.x[c - sz]);
  const S tz = (b0.x[c] - b0.x[c - sy]) - (b0.y[c] - b0.y[c - 1]);
  t0.x[c] = tx;
  t0.y[c] = ty;
  t0.z[c] = tz;
  if (tx != (S)0 || ty != (S)0 || tz != (S)0) atomicOr(nonzero, 1);
}

// ---------------------------------------------------------------------------
// Boundary hooks.

// user_loop.F90:101-164.  p: 0 xc, 1 yc1, 2 yc2, 3 zc, 4 bnorm, 5 r0,
// 6 beta_twist, 7 t0, 8 mode:
//   0  the floor/top copy as shipped (z-ghost planes only);
//   1  the footpoint-rotation driver of the commented block :122-150, with
//      its tanh(time/t0) ramp on, on the floor ghosts; top copy;
//   2  a driven loop in the boundary form of user_uniform.F90:113-142: the
//      driver's ex, ey on z <= 0 (the interior floor plane included), bz of
//      the potential field on the floor ghosts z < 0, and a conducting top
//      (ex = ey = 0, bz = the potential field).  bz of the floor plane is
//      left to Faraday's law under the driver's E (user_uniform.F90 pins it
//      too): the rotation's E has a non-zero discrete curl, so a pinned bz
//      would change div_h B of the floor cells every stage.
// One thread per (i, j) of the ghost-inclusive plane and per plane of
// k = -2, -1, 0, nz, nz+1; plane 0 is written in mode 2 only.
template <class P>
__device__ __forceinline__ P loop_bz(P bnorm, P rx, P ry1, P ry2, P rz) {
  const P rr1 = sqrt(rx * rx + ry1 * ry1 + rz * rz);
  const P rr2 = sqrt(rx * rx + ry2 * ry2 + rz * rz);
  return bnorm * (rz / (rr1 * rr1 * rr1) - rz / (rr2 * rr2 * rr2));
}

// The footpoint driver at (i, j, k): Ex, Ey and the potential-field Bz.
template <class P>
__device__ __forceinline__ void loop_driver(const double* p, P cc, int step, int i, int j, int k, P& ex, P& ey,
                                            P& bz) {
  const P xc = (P)p[0], yc1 = (P)p[1], yc2 = (P)p[2], zc = (P)p[3], bnorm = (P)p[4];
  const P r0 = (P)p[5], beta = (P)p[6], t0 = (P)p[7];
  const P xg = (P)i, yg = (P)j, zg = (P)k, hf = (P)0.5, time = (P)step;
  P rx, ry1, ry2, rz, rr1, bzl, omega, ee0;
  rx = xg + hf - xc; ry1 = yg - yc1; ry2 = yg - yc2; rz = zg - zc;
  rr1 = sqrt(rx * rx + ry1 * ry1 + rz * rz);
  bzl = loop_bz(bnorm, rx, ry1, ry2, rz);
  omega = exp(-rr1 / r0) * beta * cc / r0;
  ee0 = -omega * bzl * tanh(time / t0);
  ex = rx * ee0;
  rx = xg - xc; ry1 = yg + hf - yc1; ry2 = yg + hf - yc2; rz = zg - zc;
  rr1 = sqrt(rx * rx + ry1 * ry1 + rz * rz);
  bzl = loop_bz(bnorm, rx, ry1, ry2, rz);
  omega = exp(-rr1 / r0) * beta * cc / r0;
  ee0 = -omega * bzl * tanh(time / t0);
  ey = ry1 * ee0;
  bz = loop_bz(bnorm, xg + hf - xc, yg + hf - yc1, yg + hf - yc2, zg - zc);
}

// The conducting top of the driven loop: Bz of the potential field.
template <class P>
__device__ __forceinline__ P loop_top_bz(const double* p, int i, int j, int k) {
  const P xc = (P)p[0], yc1 = (P)p[1], yc2 = (P)p[2], zc = (P)p[3], bnorm = (P)p[4], hf = (P)0.5;
  return loop_bz(bnorm, (P)i + hf - xc, (P)j + hf - yc1, (P)j + hf - yc2, (P)k - zc);
}

template <class PR>
__global__ void k_bc_loop(Geom g, V3<typename PR::S> e, V3<typename PR::S> b, const double* p,
                          typename PR::P cc, const int* step_ptr, int step_val, int64_t lo) {
  using S = typename PR::S;
  using P = typename PR::P;
  // lo != 0 (compensated pairs): a written value's lo is 0, a copy's is the source's
  auto put = [lo](S* a, int64_t c, P v) {
    a[c] = (S)v;
    if (lo) a[c + lo] = (S)0;
  };
  auto cpy = [lo](S* a, int64_t c, int64_t src) {
    a[c] = a[src];
    if (lo) a[c + lo] = a[src + lo];
  };
  const int sx = g.nx + 2 * NG, sy = g.ny + 2 * NG;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)(2 * NG + 1) * sx * sy) return;
  const uint32_t u = (uint32_t)t;  // < 5 planes of cells: 32-bit arithmetic
  const int i = (int)(u % (uint32_t)sx) - NG;
  const int j = (int)((u / (uint32_t)sx) % (uint32_t)sy) - NG;
  const int kk = (int)(u / ((uint32_t)sx * (uint32_t)sy));
  const int k = kk < NG ? kk - NG : (kk == NG ? 0 : g.nz + kk - NG - 1);
  const int mode = (int)p[8];
  if (k == 0 && mode != 2) return;
  const int64_t c = g.at(i, j, k);
  if (k <= 0) {
    if (mode != 0) {
      P ex, ey, bz;
      loop_driver<P>(p, cc, step_ptr ? *step_ptr + step_val : step_val, i, j, k, ex, ey, bz);
      put(e.x, c, ex);
      put(e.y, c, ey);
      if (k < 0) put(b.z, c, bz);
    } else {
      const int64_t s = g.at(i, j, 0);
      cpy(e.x, c, s);
      cpy(e.y, c, s);
      cpy(b.z, c, s);
    }
  } else if (mode == 2) {
    put(b.z, c, loop_top_bz<P>(p, i, j, k));
    put(e.x, c, (P)0);
    put(e.y, c, (P)0);
  } else {
    const int64_t s = g.at(i, j, g.nz - 1);
    cpy(e.x, c, s);
    cpy(e.y, c, s);
    cpy(b.z, c, s);
  }
}

// user_uniform.F90:77-147.  p: 0 xc, 1 yc, 2 r0 (= r_L0*L0), 3 beta_twist,
// 4 t0, 5 bnorm.  Writes the planes k <= 0 (the interior plane k = 0
// included, as z_glob .le. 0 does) and the top ghost planes.
template <class PR>
__global__ void k_bc_uniform(Geom g, V3<typename PR::S> e, V3<typename PR::S> b, const double* p,
                             typename PR::P cc, const int* step_ptr, int step_val, int64_t lo) {
  using S = typename PR::S;
  using P = typename PR::P;
  auto put = [lo](S* a, int64_t c, P v) {
    a[c] = (S)v;
    if (lo) a[c + lo] = (S)0;
  };
  const int sx = g.nx + 2 * NG, sy = g.ny + 2 * NG;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)(2 * NG + 1) * sx * sy) return;
  const uint32_t u = (uint32_t)t;  // < 5 planes of cells: 32-bit arithmetic
  const int i = (int)(u % (uint32_t)sx) - NG;
  const int j = (int)((u / (uint32_t)sx) % (uint32_t)sy) - NG;
  const int kk = (int)(u / ((uint32_t)sx * (uint32_t)sy));
  const int k = kk <= NG ? kk - NG : g.nz + kk - NG - 1;
  const int64_t c = g.at(i, j, k);
  const int step = step_ptr ? *step_ptr + step_val : step_val;
  const P xc = (P)p[0], yc = (P)p[1], r0 = (P)p[2], beta = (P)p[3], t0 = (P)p[4], bnorm = (P)p[5];
  const P xg = (P)i, yg = (P)j, time = (P)step, hf = (P)0.5, tiny = (P)1e-7, one = (P)1.0;
  P rx, ry, rr, omega, ee0;
  if (k <= 0) {
    rx = xg + hf - xc; ry = yg - yc;
    rr = sqrt(rx * rx + ry * ry + tiny);
    omega = exp(-rr / r0) * beta * cc / r0;
    ee0 = -omega * bnorm * tanh(time / t0);
    put(e.x, c, rx * ee0);
    rx = xg - xc; ry = yg + hf - yc;
    rr = sqrt(rx * rx + ry * ry + tiny);
    omega = exp(-rr / r0) * beta * cc / r0;
    ee0 = -omega * bnorm * tanh(time / t0);
    put(e.y, c, ry * ee0);
    rx = xg + hf - xc; ry = yg + hf - yc;
    rr = sqrt(rx * rx + ry * ry + tiny);
    put(b.z, c, bnorm / (one + (rr / r0)));
  } else {
    rx = xg + hf - xc; ry = yg + hf - yc;
    rr = sqrt(rx * rx + ry * ry + tiny);
    put(b.z, c, bnorm / (one + (rr / r0)));
    put(e.x, c, (P)0);
    put(e.y, c, (P)0);
  }
}

// user_psr.F90:345-374 getDipole (p: 0 psr_radius, 1 psr_angle [rad],
// 2 psr_omega, 3 xc_g, 4 yc_g, 5 zc_g)
template <class P>
__device__ __forceinline__ void dipole(const double* p, int step, P xg, P yg, P zg, P& obx, P& oby,
                                       P& obz) {
  const P R = (P)p[0], ang = (P)p[1], om = (P)p[2], xc = (P)p[3], yc = (P)p[4], zc = (P)p[5];
  const P phase = om * (P)step;
  P nx = xg - xc, ny = yg - yc, nz = zg - zc;
  P rr = (P)1.0 / sqrt(nx * nx + ny * ny + nz * nz);
  nx = nx * rr; ny = ny * rr; nz = nz * rr;
  rr = rr * rr * rr;
  const P r3 = R * R * R, off = (P)0.0;
  const P mux = r3 * sin(ang) * cos(phase + off);
  const P muy = r3 * sin(ang) * sin(phase + off);
  const P muz = r3 * cos(ang);
  const P mdn = mux * nx + muy * ny + muz * nz;
  obx = ((P)3.0 * nx * mdn - mux) * rr;
  oby = ((P)3.0 * ny * mdn - muy) * rr;
  obz = ((P)3.0 * nz * mdn - muz) * rr;
}

// user_psr.F90:376-383 shape
template <class P>
__device__ __forceinline__ P psr_shape(P rad, P rad0) {
  const P del = (P)1.0;
  return (P)0.5 * ((P)1.0 - tanh((rad - rad0) / del));
}

// user_psr.F90:147-302: conducting-sphere blend for interior cells inside
// the supersphere r < 2 R, evaluated on the box [lo, lo+n) and written to
// the staging arrays `nw` (6 components, box layout); k_bc_psr_commit then
// copies them back (:305-325), so every cell reads pre-BC neighbours.
template <class PR>
__global__ void k_bc_psr_eval(Geom g, V3<const typename PR::S> e, V3<const typename PR::S> b,
                              const double* p, typename PR::P ccinv, int3 lo, int3 n,
                              typename PR::S* nw, const int* step_ptr, int step_val) {
  using S = typename PR::S;
  using T = typename PR::T;
  using P = typename PR::P;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t nbox = (int64_t)n.x * n.y * n.z;
  if (t >= nbox) return;
  const uint32_t u = (uint32_t)t;  // the bounding box of r < 2R: 32-bit arithmetic
  const int i = lo.x + (int)(u % (uint32_t)n.x);
  const int j = lo.y + (int)((u / (uint32_t)n.x) % (uint32_t)n.y);
  const int k = lo.z + (int)(u / ((uint32_t)n.x * (uint32_t)n.y));
  const P R = (P)p[0], om = (P)p[2], xc = (P)p[3], yc = (P)p[4], zc = (P)p[5];
  const P ssq = (R * (P)2) * (R * (P)2);
  const P fi = (P)i, fj = (P)j, fk = (P)k;
  if (!((fi - xc) * (fi - xc) + (fj - yc) * (fj - yc) + (fk - zc) * (fk - zc) < ssq)) return;
  const int step = step_ptr ? *step_ptr + step_val : step_val;
  const int64_t c = g.at(i, j, k), sy = g.px, sz = g.plane;
#define A(arr, a, bb, d) ((P)(T)arr[c + (a) + (bb)*sy + (d)*sz])
  const P shift_B = (P)4.0, shift_E = (P)0.0, one = (P)1.0, hf = (P)0.5, qq = (P)0.25;
  const P sEpar = (P)0.5, sEperp = (P)0.25, sBperp = (P)0.5, sBpar = (P)0.25;
  P rx, ry, rz, rr_sqr, s, scale, bdx, bdy, bdz, bidr, bddr, v;
  P out[6];
  // B_x (:157-174)
  rx = fi - xc; ry = fj + hf - yc; rz = fk + hf - zc;
  rr_sqr = rx * rx + ry * ry + rz * rz;
  bidr = A(b.x, 0, 0, 0) * rx + qq * (A(b.y, 0, 0, 0) + A(b.y, 0, 1, 0) + A(b.y, -1, 0, 0) + A(b.y, -1, 1, 0)) * ry +
         qq * (A(b.z, 0, 0, 0) + A(b.z, 0, 0, 1) + A(b.z, -1, 0, 0) + A(b.z, -1, 0, 1)) * rz;
  dipole<P>(p, step, rx + xc, ry + yc, rz + zc, bdx, bdy, bdz);
  bddr = bdx * rx + bdy * ry + bdz * rz;
  scale = one / sBperp;
  s = psr_shape<P>(sqrt(rr_sqr) * scale, (R - shift_B) * scale);
  v = (bidr - bddr) * (rx / rr_sqr) * (one - s);
  scale = one / sBpar;
  s = psr_shape<P>(sqrt(rr_sqr) * scale, (R - shift_B) * scale);
  out[3] = v + bdx + ((A(b.x, 0, 0, 0) - bidr * rx / rr_sqr) - (bdx - bddr * rx / rr_sqr)) * (one - s);
  // B_y (:177-194)
  rx = fi + hf - xc; ry = fj - yc; rz = fk + hf - zc;
  rr_sqr = rx * rx + ry * ry + rz * rz;
  bidr = qq * (A(b.x, 0, 0, 0) + A(b.x, 1, 0, 0) + A(b.x, 0, -1, 0) + A(b.x, 1, -1, 0)) * rx + A(b.y, 0, 0, 0) * ry +
         qq * (A(b.z, 0, 0, 0) + A(b.z, 0, 0, 1) + A(b.z, 0, -1, 0) + A(b.z, 0, -1, 1)) * rz;
  dipole<P>(p, step, rx + xc, ry + yc, rz + zc, bdx, bdy, bdz);
  bddr = bdx * rx + bdy * ry + bdz * rz;
  scale = one / sBperp;
  s = psr_shape<P>(sqrt(rr_sqr) * scale, (R - shift_B) * scale);
  v = (bidr - bddr) * (ry / rr_sqr) * (one - s);
  scale = one / sBpar;
  s = psr_shape<P>(sqrt(rr_sqr) * scale, (R - shift_B) * scale);
  out[4] = v + bdy + ((A(b.y, 0, 0, 0) - bidr * ry / rr_sqr) - (bdy - bddr * ry / rr_sqr)) * (one - s);
  // B_z (:197-214)
  rx = fi + hf - xc; ry = fj + hf - yc; rz = fk - zc;
  rr_sqr = rx * rx + ry * ry + rz * rz;
  bidr = qq * (A(b.x, 0, 0, 0) + A(b.x, 1, 0, 0) + A(b.x, 0, 0, -1) + A(b.x, 1, 0, -1)) * rx +
         qq * (A(b.y, 0, 0, 0) + A(b.y, 0, 1, 0) + A(b.y, 0, 0, -1) + A(b.y, 0, 1, -1)) * ry + A(b.z, 0, 0, 0) * rz;
  dipole<P>(p, step, rx + xc, ry + yc, rz + zc, bdx, bdy, bdz);
  bddr = bdx * rx + bdy * ry + bdz * rz;
  scale = one / sBperp;
  s = psr_shape<P>(sqrt(rr_sqr) * scale, (R - shift_B) * scale);
  v = (bidr - bddr) * (rz / rr_sqr) * (one - s);
  scale = one / sBpar;
  s = psr_shape<P>(sqrt(rr_sqr) * scale, (R - shift_B) * scale);
  out[5] = v + bdz + ((A(b.z, 0, 0, 0) - bidr * rz / rr_sqr) - (bdz - bddr * rz / rr_sqr)) * (one - s);

  P vx, vy, vz, edx, edy, edz, eidr, eddr;
  // E_x (:219-243)
  rx = fi + hf - xc; ry = fj - yc; rz = fk - zc;
  rr_sqr = rx * rx + ry * ry + rz * rz;
  eidr = A(e.x, 0, 0, 0) * rx + qq * (A(e.y, 0, 0, 0) + A(e.y, 1, 0, 0) + A(e.y, 0, -1, 0) + A(e.y, 1, -1, 0)) * ry +
         qq * (A(e.z, 0, 0, 0) + A(e.z, 1, 0, 0) + A(e.z, 0, 0, -1) + A(e.z, 1, 0, -1)) * rz;
  dipole<P>(p, step, rx + xc, ry + yc, rz + zc, bdx, bdy, bdz);
  vx = -om * ry; vy = om * rx; vz = (P)0.0;
  edx = -(vy * bdz - vz * bdy) * ccinv;
  edy = (vx * bdz - vz * bdx) * ccinv;
  edz = -(vx * bdy - vy * bdx) * ccinv;
  eddr = edx * rx + edy * ry + edz * rz;
  scale = one / sEpar;
  s = psr_shape<P>(sqrt(rr_sqr) * scale, (R - shift_E) * scale);
  v = edx + ((A(e.x, 0, 0, 0) - eidr * rx / rr_sqr) - (edx - eddr * rx / rr_sqr)) * (one - s);
  scale = one / sEperp;
  s = psr_shape<P>(sqrt(rr_sqr) * scale, (R - shift_E) * scale);
  out[0] = v + (eidr - eddr) * (rx / rr_sqr) * (one - s);
  // E_y (:246-270)
  rx = fi - xc; ry = fj + hf - yc; rz = fk - zc;
  rr_sqr = rx * rx + ry * ry + rz * rz;
  eidr = qq * (A(e.x, 0, 0, 0) + A(e.x, 0, 1, 0) + A(e.x, -1, 0, 0) + A(e.x, -1, 1, 0)) * rx + A(e.y, 0, 0, 0) * ry +
         qq * (A(e.z, 0, 0, 0) + A(e.z, 0, 1, 0) + A(e.z, 0, 0, -1) + A(e.z, 0, 1, -1)) * rz;
  dipole<P>(p, step, rx + xc, ry + yc, rz + zc, bdx, bdy, bdz);
  vx = -om * ry; vy = om * rx; vz = (P)0.0;
  edx = -(vy * bdz - vz * bdy) * ccinv;
  edy = (vx * bdz - vz * bdx) * ccinv;
  edz = -(vx * bdy - vy * bdx) * ccinv;
  eddr = edx * rx + edy * ry + edz * rz;
  scale = one / sEpar;
  s = psr_shape<P>(sqrt(rr_sqr) * scale, (R - shift_E) * scale);
  v = edy + ((A(e.y, 0, 0, 0) - eidr * ry / rr_sqr) - (edy - eddr * ry / rr_sqr)) * (one - s);
  scale = one / sEperp;
  s = psr_shape<P>(sqrt(rr_sqr) * scale, (R - shift_E) * scale);
  out[1] = v + (eidr - eddr) * (ry / rr_sqr) * (one - s);
  // E_z (:273-297)
  rx = fi - xc; ry = fj - yc; rz = fk + hf - zc;
  rr_sqr = rx * rx + ry * ry + rz * rz;
  eidr = qq * (A(e.x, 0, 0, 0) + A(e.x, 0, 0, 1) + A(e.x, -1, 0, 0) + A(e.x, -1, 0, 1)) * rx +
         qq * (A(e.y, 0, 0, 0) + A(e.y, 0, -1, 0) + A(e.y, 0, 0, 1) + A(e.y, 0, -1, 1)) * ry + A(e.z, 0, 0, 0) * rz;
  dipole<P>(p, step, rx + xc, ry + yc, rz + zc, bdx, bdy, bdz);
  vx = -om * ry; vy = om * rx; vz = (P)0.0;
  edx = -(vy * bdz - vz * bdy) * ccinv;
  edy = (vx * bdz - vz * bdx) * ccinv;
  edz = -(vx * bdy - vy * bdx) * ccinv;
  eddr = edx * rx + edy * ry + edz * rz;
  scale = one / sEpar;
  s = psr_shape<P>(sqrt(rr_sqr) * scale, (R - shift_E) * scale);
  v = edz + ((A(e.z, 0, 0, 0) - eidr * rz / rr_sqr) - (edz - eddr * rz / rr_sqr)) * (one - s);
  scale = one / sEperp;
  s = psr_shape<P>(sqrt(rr_sqr) * scale, (R - shift_E) * scale);
  out[2] = v + (eidr - eddr) * (rz / rr_sqr) * (one - s);
#undef A
  for (int f = 0; f < 6; ++f) nw[f * nbox + t] = (S)out[f];
}

template <class PR>
__global__ void k_bc_psr_commit(Geom g, V3<typename PR::S> e, V3<typename PR::S> b, const double* p,
                                int3 lo, int3 n, const typename PR::S* nw, int64_t pair) {
  using P = typename PR::P;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t nbox = (int64_t)n.x * n.y * n.z;
  if (t >= nbox) return;
  const uint32_t u = (uint32_t)t;  // the bounding box of r < 2R: 32-bit arithmetic
  const int i = lo.x + (int)(u % (uint32_t)n.x);
  const int j = lo.y + (int)((u / (uint32_t)n.x) % (uint32_t)n.y);
  const int k = lo.z + (int)(u / ((uint32_t)n.x * (uint32_t)n.y));
  const P R = (P)p[0], xc = (P)p[3], yc = (P)p[4], zc = (P)p[5];
  const P ssq = (R * (P)2) * (R * (P)2);
  const P fi = (P)i, fj = (P)j, fk = (P)k;
  if (!((fi - xc) * (fi - xc) + (fj - yc) * (fj - yc) + (fk - zc) * (fk - zc) < ssq)) return;
  const int64_t c = g.at(i, j, k);
  typename PR::S* dst[6] = {e.x, e.y, e.z, b.x, b.y, b.z};
#pragma unroll
  for (int m = 0; m < 6; ++m) {
    dst[m][c] = nw[m * nbox + t];
    if (pair) dst[m][c + pair] = (typename PR::S)0;  // compensated pairs: the blend's lo is 0
  }
}

// ---------------------------------------------------------------------------
// FP64 diagnostics: per-block partial sums of E^2, B^2, (div_h B)^2 and the
// block maximum of |div_h B| (wave64 shuffles, then LDS across waves); a
// single-block pass combines the partials in a fixed order, so the result
// does not depend on scheduling.
constexpr int RED_THREADS = 256;

__device__ __forceinline__ void wave_reduce(double& a, double& b, double& c, double& m) {
  for (int off = 32; off > 0; off >>= 1) {
    a += __shfl_down(a, off, 64);
    b += __shfl_down(b, off, 64);
    c += __shfl_down(c, off, 64);
    m = fmax(m, __shfl_down(m, off, 64));
  }
}

template <class S>
__global__ void __launch_bounds__(RED_THREADS)
    k_reduce_partial(Geom g, V3<const S> e, V3<const S> b, double* part, int64_t lo) {
  const int64_t ncell = (int64_t)g.nx * g.ny * g.nz;
  double se = 0.0, sb = 0.0, sd = 0.0, md = 0.0;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < ncell;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(t % g.nx);
    const int j = (int)((t / g.nx) % g.ny);
    const int k = (int)(t / ((int64_t)g.nx * g.ny));
    const int64_t c = g.at(i, j, k);
    // lo != 0: the compensated pairs, hi + lo (exact in FP64)
    auto v = [&](const S* a, int64_t o) { return lo ? (double)a[o] + (double)a[o + lo] : (double)a[o]; };
    const double ax = v(e.x, c), ay = v(e.y, c), az = v(e.z, c);
    const double bx = v(b.x, c), by = v(b.y, c), bz = v(b.z, c);
    se += ax * ax + ay * ay + az * az;
    sb += bx * bx + by * by + bz * bz;
    const double dv = (v(b.x, c + 1) - bx) + (v(b.y, c + g.px) - by) + (v(b.z, c + g.plane) - bz);
    sd += dv * dv;
    md = fmax(md, fabs(dv));
  }
  wave_reduce(se, sb, sd, md);
  __shared__ double sh[4][RED_THREADS / 64];
  const int w = threadIdx.x / 64, l = threadIdx.x % 64;
  if (l == 0) {
    sh[0][w] = se; sh[1][w] = sb; sh[2][w] = sd; sh[3][w] = md;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = 0, bb = 0, cc = 0, m = 0;
    for (int q = 0; q < RED_THREADS / 64; ++q) {
      a += sh[0][q]; bb += sh[1][q]; cc += sh[2][q]; m = fmax(m, sh[3][q]);
    }
    part[4 * blockIdx.x + 0] = a;
    part[4 * blockIdx.x + 1] = bb;
    part[4 * blockIdx.x + 2] = cc;
    part[4 * blockIdx.x + 3] = m;
  }
}

__global__ void __launch_bounds__(64) k_reduce_final(const double* part, int nblocks, double* out,
                                                     double ncell) {
  double a = 0, b = 0, c = 0, m = 0;
  for (int q = threadIdx.x; q < nblocks; q += 64) {
    a += part[4 * q]; b += part[4 * q + 1]; c += part[4 * q + 2]; m = fmax(m, part[4 * q + 3]);
  }
  wave_reduce(a, b, c, m);
  if (threadIdx.x == 0) {
    out[0] = 0.5 * a;
    out[1] = 0.5 * b;
    out[2] = m;
    out[3] = sqrt(c / ncell);
  }
}

__global__ void k_bump_step(int* step, int by) {
  if (threadIdx.x == 0) *step += by;
}

// sets a device counter in stream order (a host value copied from the stack
// would not be)
__global__ void k_set_int(int* p, int v) {
  if (threadIdx.x == 0) *p = v;
}

template <class S>
__global__ void k_copy3(int64_t n, V3<const S> a, V3<S> b) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  b.x[t] = a.x[t];
  b.y[t] = a.y[t];
  b.z[t] = a.z[t];
}

}  // namespace tff
