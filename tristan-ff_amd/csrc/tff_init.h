// tff_init.h -- initial conditions generated in HBM (the user files'
// userInitFields, and a seeded Alfvenic spectrum), so that large boxes need
// no host-built arrays.  Every value is computed in FP64 and stored in the
// handle's precision.  One thread per cell of the ghost-inclusive box.
#pragma once
#include "tff_kernels.h"

namespace tff {

// ghost-inclusive cell t -> (i, j, k), i fastest
__device__ __forceinline__ void box_cell(const Geom& g, int64_t t, int& i, int& j, int& k) {
  const int64_t sx = g.nx + 2 * NG, sy = g.ny + 2 * NG;
  i = (int)(t % sx) - NG;
  j = (int)((t / sx) % sy) - NG;
  k = (int)(t / (sx * sy)) - NG;
}

__device__ __forceinline__ bool interior_cell(const Geom& g, int i, int j, int k) {
  return i >= 0 && i < g.nx && j >= 0 && j < g.ny && k >= 0 && k < g.nz;
}

// Stores the nine values of cell c; lo != 0 (a compensated handle) also
// stores each value's remainder x - fl(x) `lo` elements after it.
template <class S>
__device__ __forceinline__ void store6(const Geom& g, int64_t c, V3<S> e, V3<S> b, V3<S> b0, double ex, double ey,
                                       double ez, double bx, double by, double bz, double b0x, double b0y,
                                       double b0z, int64_t lo) {
  S* dst[9] = {e.x, e.y, e.z, b.x, b.y, b.z, b0.x, b0.y, b0.z};
  const double val[9] = {ex, ey, ez, bx, by, bz, b0x, b0y, b0z};
#pragma unroll
  for (int q = 0; q < 9; ++q) {
    const S h = (S)val[q];
    dst[q][c] = h;
    if (lo) dst[q][c + lo] = (S)(val[q] - (double)h);
  }
}

// user_loop.F90:44-93: two opposite sources at (xc, yc1|yc2, zc); every cell
// including the ghosts; E = 0, B0 = B.  p: xc, yc1, yc2, zc, bnorm.
template <class S>
__global__ void k_init_loop(Geom g, V3<S> e, V3<S> b, V3<S> b0, const double* p, int64_t ncell, int64_t lo) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ncell) return;
  int i, j, k;
  box_cell(g, t, i, j, k);
  const double xc = p[0], yc1 = p[1], yc2 = p[2], zc = p[3], bnorm = p[4];
  const double x = i, y = j, z = k;
  auto field = [&](double rx, double ry1, double ry2, double rz, double n1, double n2) {
    const double rr1 = sqrt(rx * rx + ry1 * ry1 + rz * rz);
    const double rr2 = sqrt(rx * rx + ry2 * ry2 + rz * rz);
    return bnorm * (n1 / (rr1 * rr1 * rr1) - n2 / (rr2 * rr2 * rr2));
  };
  double rx = x - xc, ry1 = y + 0.5 - yc1, ry2 = y + 0.5 - yc2, rz = z + 0.5 - zc;
  const double bx = field(rx, ry1, ry2, rz, rx, rx);
  rx = x + 0.5 - xc; ry1 = y - yc1; ry2 = y - yc2; rz = z + 0.5 - zc;
  const double by = field(rx, ry1, ry2, rz, ry1, ry2);
  rx = x + 0.5 - xc; ry1 = y + 0.5 - yc1; ry2 = y + 0.5 - yc2; rz = z - zc;
  const double bz = field(rx, ry1, ry2, rz, rz, rz);
  store6(g, g.at(i, j, k), e, b, b0, 0.0, 0.0, 0.0, bx, by, bz, bx, by, bz, lo);
}

// user_psr.F90:39-69: the dipole of radius R at (xc, yc, zc) on the
// interior cells at phase 0, zero in the ghosts; E = 0, B0 = 0.
// p: R, angle [rad], omega (unused), xc, yc, zc.
template <class S>
__global__ void k_init_psr(Geom g, V3<S> e, V3<S> b, V3<S> b0, const double* p, int64_t ncell, int64_t lo) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ncell) return;
  int i, j, k;
  box_cell(g, t, i, j, k);
  double bx = 0.0, by = 0.0, bz = 0.0;
  if (interior_cell(g, i, j, k)) {
    double ox, oy, oz;
    dipole<double>(p, 0, (double)i, j + 0.5, k + 0.5, bx, oy, oz);
    dipole<double>(p, 0, i + 0.5, (double)j, k + 0.5, ox, by, oz);
    dipole<double>(p, 0, i + 0.5, j + 0.5, (double)k, ox, oy, bz);
  }
  store6(g, g.at(i, j, k), e, b, b0, 0.0, 0.0, 0.0, bx, by, bz, 0.0, 0.0, 0.0, lo);
}

// user_uniform.F90:43-69: bz = bnorm / (1 + r/r0) about the vertical axis
// through (xc, yc), every cell including the ghosts; E = 0, B0 = B.
// p: xc, yc, r0, bnorm.
template <class S>
__global__ void k_init_uniform(Geom g, V3<S> e, V3<S> b, V3<S> b0, const double* p, int64_t ncell,
                               int64_t lo) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ncell) return;
  int i, j, k;
  box_cell(g, t, i, j, k);
  const double rx = i + 0.5 - p[0], ry = j + 0.5 - p[1];
  const double rr = sqrt(rx * rx + ry * ry + 1e-7);  // TINY, src/defs.F90:4
  const double bz = p[3] / (1.0 + rr / p[2]);
  store6(g, g.at(i, j, k), e, b, b0, 0.0, 0.0, 0.0, 0.0, 0.0, bz, 0.0, 0.0, bz, lo);
}

// A sum of Fourier modes on the periodic box: component q of cell (i, j, k)
// = background[q] + sum_m Re[c_{m,q} exp(2 pi i (kx i / nx + ky j / ny +
// kz k / nz))], for q = ex ey ez bx by bz, on the interior cells (ghosts
// zero; a periodic box's ghosts are never read before the first fill);
// B0 = the background B everywhere.  A mode is 15 doubles: kx ky kz
// (integers), then (re, im) of c for the six components.
constexpr int SPEC_MODE = 15;
template <class S>
__global__ void k_init_spectrum(Geom g, V3<S> e, V3<S> b, V3<S> b0, const double* modes, int nmodes,
                                const double* bg, int64_t ncell, int64_t lo) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ncell) return;
  int i, j, k;
  box_cell(g, t, i, j, k);
  double v[6] = {bg[0], bg[1], bg[2], bg[3], bg[4], bg[5]};
  if (interior_cell(g, i, j, k)) {
    const double tau = 6.283185307179586476925286766559;
    // k.x / n reduced per axis in integers first: the phase's fraction of
    // a turn is exact to one rounding whatever the size of k * i
    // (tff_init_spectrum checks |k| (n + 4) < 2^31, so k * x fits 32 bits)
    auto turn = [](double kq, int x, int nq) {
      const int r = ((int)kq * x) % nq;
      return (double)(r < 0 ? r + nq : r) / nq;
    };
    for (int m = 0; m < nmodes; ++m) {
      const double* c = modes + (size_t)m * SPEC_MODE;
      const double ph = tau * (turn(c[0], i, g.nx) + turn(c[1], j, g.ny) + turn(c[2], k, g.nz));
      double sn, cs;
      sincos(ph, &sn, &cs);
#pragma unroll
      for (int q = 0; q < 6; ++q) v[q] += c[3 + 2 * q] * cs - c[4 + 2 * q] * sn;
    }
  } else {
    for (int q = 0; q < 6; ++q) v[q] = 0.0;
  }
  store6(g, g.at(i, j, k), e, b, b0, v[0], v[1], v[2], v[3], v[4], v[5], bg[3], bg[4], bg[5], lo);
}

}  // namespace tff
