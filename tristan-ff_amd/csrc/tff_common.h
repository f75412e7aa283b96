// tff_common.h -- device-side building blocks shared by the force-free
// kernels: mesh geometry, precision policies and the per-cell arithmetic of
// the reference's field_solver.F90 routines.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tff {

constexpr int NG = 2;  // NGHOST, src/CMakeLists.txt:19

// |Bbar|^2 below this fraction of the cell's own |B|^2 counts as near a null
constexpr float NEAR_NULL = 1e-4f;
// The fast path's cleanEpar rescales a B interpolant whose |Bbar|^2 is
// non-zero and below this by 2^64 before the projection: the hardware
// reciprocal overflows near 1e-38.  (An exact zero keeps the guard's
// result; a zero that cancels a non-zero B falls under NEAR_NULL.)
constexpr float TINY_B2 = 1e-30f;

// Device layout of one field component: x fastest, every x-row padded to
// `px` elements so that the first interior cell of each row is 128-byte
// aligned; element (i,j,k), i,j,k in [-NG, n-1+NG], lives at
// org + i + j*px + k*plane.
struct Geom {
  int nx, ny, nz;
  int px;
  int64_t plane;
  int64_t org;
  int64_t cs;  // elements between the x, y and z components of one field buffer
  __host__ __device__ int64_t at(int i, int j, int k) const {
    return org + i + (int64_t)j * px + (int64_t)k * plane;
  }
};

template <class S>
struct V3 {
  S* x;
  S* y;
  S* z;
};

// Precision policies: S = storage, T = stencil arithmetic (differences and
// interpolation sums), Q = normalisation by |B|^2, E.B projection and
// limiter, P = stage accumulation, the curl terms and boundary geometry.
// With null_fp64 (the FP32-bulk policy), arithmetic contracts into fused
// multiply-adds, divides through the hardware reciprocal and updates in the
// scale-exact form of update_cell; a cell whose interpolated |B|^2 is small
// next to its own |B|^2 (a magnetic null, where the interpolation cancels)
// redoes the interpolation, projection and currents in FP64.  TFF_BULK_P
// selects FP64 accumulation for that policy instead (a build option for
// comparisons; FP32 with the scale-exact update is as accurate and
// faster).
struct PrecF64 {
  using S = double;
  using T = double;
  using Q = double;
  using P = double;
  static constexpr bool null_fp64 = false;
};
struct PrecF32 {
  using S = float;
  using T = float;
  using Q = float;
  using P = float;
  static constexpr bool null_fp64 = false;
};
#ifndef TFF_BULK_P
#define TFF_BULK_P float
#endif
struct PrecF32Bulk {
  using S = float;
  using T = float;
  using Q = float;
  using P = TFF_BULK_P;
  static constexpr bool null_fp64 = true;
};

// Parameters of rk3Update's absorbing layer (field_solver.F90:198-269).
template <class P>
struct Absorb {
  int on;
  int fixed;  // 0: jg = i + y0 and delta unclamped (as coded); 1: fixed
  P xc, yc, zc, rabsorb, buff;
};

// Arithmetic helpers.  FAST (the FP32-bulk policy) contracts a*b + c into one
// fused multiply-add and divides through the hardware reciprocal (1 ulp);
// otherwise every operation rounds separately, as the reference does.
template <bool FAST, class Q>
__device__ __forceinline__ Q mad_(Q a, Q b, Q c) {
  if constexpr (FAST) return __builtin_fmaf(a, b, c);
  return a * b + c;
}
template <bool FAST>
__device__ __forceinline__ double mad_(double a, double b, double c) {
  if constexpr (FAST) return __builtin_fma(a, b, c);
  return a * b + c;
}
template <bool FAST, class Q>
__device__ __forceinline__ Q div_(Q a, Q b) {
  if constexpr (FAST) return a * __builtin_amdgcn_rcpf(b);
  return a / b;
}
template <bool FAST>
__device__ __forceinline__ double div_(double a, double b) {
  return a / b;
}
// |a|^2 of a 3-vector in the operation order x, y, z
template <bool FAST, class Q>
__device__ __forceinline__ Q sq3(Q x, Q y, Q z) {
  if constexpr (FAST) return mad_<FAST>(z, z, mad_<FAST>(y, y, x * x));
  return x * x + y * y + z * z;
}

// Drift current component (field_solver.F90:103-111 and cyclic forms):
// the limiter rescales the two interpolated E components transverse to
// the current's own axis when |E| > |B|, then
// j = CC / (|B|^2 + TINY) * (rho * (Ea*Bb - Ba*Eb)).
// `self` is the E component along the current's axis (never rescaled).
// `bmag` is |B|^2 at the point, summed in the order x, y, z.
template <class P, bool FAST = false>
__device__ __forceinline__ P drift_current(P cc, P intrho, P self, P ea, P eb, P bs, P ba, P bb, P bmag,
                                           int comp, P bref, bool& near) {
  // comp 0 (x): self=Ex, ea=Ey, eb=Ez, bs=Bx, ba=By, bb=Bz
  // comp 1 (y): self=Ey, ea=Ex, eb=Ez, bs=By, ba=Bx, bb=Bz
  // comp 2 (z): self=Ez, ea=Ex, eb=Ey, bs=Bz, ba=Bx, bb=By
  // (sums of squares in the order x, y, z)
  P emag;
  if (comp == 0)
    emag = sq3<FAST>(self, ea, eb);
  else if (comp == 1)
    emag = sq3<FAST>(ea, self, eb);
  else
    emag = sq3<FAST>(ea, eb, self);
  near |= bmag < (P)NEAR_NULL * bref;
  if (emag > bmag) {
    // IEEE division even on the fast path: this branch is rare, and the
    // hardware reciprocal of an |E|^2 below 2^-128 overflows to +inf, which
    // turns bmag = 0 (a zeroed region) into 0 * inf = NaN
    P corr = sqrt(bmag / emag);
    ea = ea * corr;
    eb = eb * corr;
  }
  const P tiny = (P)1e-7;  // src/defs.F90:4
  P cross;
  if (comp == 0)
    cross = mad_<FAST>(ea, bb, -(ba * eb));  // intey*intbz - intby*intez
  else if (comp == 1)
    cross = mad_<FAST>(eb, ba, -(ea * bb));  // intez*intbx - intex*intbz
  else
    cross = mad_<FAST>(ea, bb, -(ba * eb));  // intex*intby - intbx*intey
  if constexpr (FAST) return div_<FAST>(cc * (intrho * cross), bmag + tiny);
  return cc / (bmag + tiny) * (intrho * cross);
}

}  // namespace tff
