// tff_common.h -- device-side building blocks shared by the force-free
// kernels: mesh geometry, precision policies and the per-cell arithmetic of
// the reference's field_solver.F90 routines.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tff {

constexpr int NG = 2;  // NGHOST, src/CMakeLists.txt:19

// Device layout of one field component: x fastest, every x-row padded to
// `px` elements so that the first interior cell of each row is 128-byte
// aligned; element (i,j,k), i,j,k in [-NG, n-1+NG], lives at
// org + i + j*px + k*plane.
struct Geom {
  int nx, ny, nz;
  int px;
  int64_t plane;
  int64_t org;
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
// interpolation sums), P = normalisation by |B|^2, E.B projection, limiter,
// stage accumulation and boundary geometry.
struct PrecF64 {
  using S = double;
  using T = double;
  using P = double;
};
struct PrecF32 {
  using S = float;
  using T = float;
  using P = float;
};
struct PrecF32Bulk {
  using S = float;
  using T = float;
  using P = double;
};

// Parameters of rk3Update's absorbing layer (field_solver.F90:198-269).
template <class P>
struct Absorb {
  int on;
  int fixed;  // 0: jg = i + y0 and delta unclamped (as coded); 1: fixed
  P xc, yc, zc, rabsorb, buff;
};

// Drift current component (field_solver.F90:103-111 and cyclic forms):
// the limiter rescales the two interpolated E components transverse to
// the current's own axis when |E| > |B|, then
// j = CC / (|B|^2 + TINY) * (rho * (Ea*Bb - Ba*Eb)).
// `self` is the E component along the current's axis (never rescaled).
template <class P>
__device__ __forceinline__ P drift_current(P cc, P intrho, P self, P ea, P eb, P bs, P ba, P bb,
                                           int comp) {
  // comp 0 (x): self=Ex, ea=Ey, eb=Ez, bs=Bx, ba=By, bb=Bz
  P bmag, emag;
  if (comp == 0) {
    bmag = bs * bs + ba * ba + bb * bb;
    emag = self * self + ea * ea + eb * eb;
  } else if (comp == 1) {
    // ordering of the sums follows x,y,z: here self=Ey, ea=Ex, eb=Ez,
    // bs=By, ba=Bx, bb=Bz
    bmag = ba * ba + bs * bs + bb * bb;
    emag = ea * ea + self * self + eb * eb;
  } else {
    // self=Ez, ea=Ex, eb=Ey, bs=Bz, ba=Bx, bb=By
    bmag = ba * ba + bb * bb + bs * bs;
    emag = ea * ea + eb * eb + self * self;
  }
  if (emag > bmag) {
    P corr = sqrt(bmag / emag);
    ea = ea * corr;
    eb = eb * corr;
  }
  const P tiny = (P)1e-7;  // src/defs.F90:4
  P cross;
  if (comp == 0)
    cross = ea * bb - ba * eb;  // intey*intbz - intby*intez
  else if (comp == 1)
    cross = eb * ba - ea * bb;  // intez*intbx - intex*intbz
  else
    cross = ea * bb - ba * eb;  // intex*intby - intbx*intey
  return cc / (bmag + tiny) * (intrho * cross);
}

}  // namespace tff
