// tff_cell.h -- per-cell arithmetic of the force-free step, written once and
// shared by every kernel (global-memory and LDS-tiled forms), so that all
// paths evaluate the same expression trees in the same order.
//
// Field access goes through accessors f(a, b, d) returning the value at
// offset (a, b, d) from the cell, already converted to the stencil type T.
#pragma once
#include "tff_common.h"

namespace tff {

// B interpolated to the three E points and |B|^2 there: index 0 is the Ex
// point (i+1/2, j, k), 1 the Ey point, 2 the Ez point.  Bx at the Ex point
// is the 8-point average over i..i+1, j-1..j, k-1..k, By and Bz the 2-point
// averages along z and y (field_solver.F90:96-101), and cyclically for the
// other points (:119-124, :142-147).  cleanEpar interpolates B with the same
// expressions (:288-328), so one evaluation serves both.  Sums in T2,
// stored in Q2; |B|^2 in the order x, y, z.
template <class Q2>
struct BBar {
  Q2 x[3], y[3], z[3], m[3];
};

// The B interpolants at E point P alone (the expressions bbar uses).
template <int P, class T2, class BXa, class BYa, class BZa>
__device__ __forceinline__ void bbar_point(const BXa& BXF, const BYa& BYF, const BZa& BZF, T2& bx, T2& by,
                                           T2& bz) {
  const T2 h = (T2)0.5, e = (T2)0.125;
  if constexpr (P == 0) {
    bx = e * ((T2)BXF(0, 0, 0) + (T2)BXF(0, -1, 0) + (T2)BXF(1, -1, 0) + (T2)BXF(1, 0, 0) +
              (T2)BXF(0, 0, -1) + (T2)BXF(0, -1, -1) + (T2)BXF(1, -1, -1) + (T2)BXF(1, 0, -1));
    by = h * ((T2)BYF(0, 0, 0) + (T2)BYF(0, 0, -1));
    bz = h * ((T2)BZF(0, 0, 0) + (T2)BZF(0, -1, 0));
  } else if constexpr (P == 1) {
    bx = h * ((T2)BXF(0, 0, 0) + (T2)BXF(0, 0, -1));
    by = e * ((T2)BYF(0, 0, 0) + (T2)BYF(-1, 0, 0) + (T2)BYF(-1, 1, 0) + (T2)BYF(0, 1, 0) +
              (T2)BYF(0, 0, -1) + (T2)BYF(-1, 0, -1) + (T2)BYF(-1, 1, -1) + (T2)BYF(0, 1, -1));
    bz = h * ((T2)BZF(0, 0, 0) + (T2)BZF(-1, 0, 0));
  } else {
    bx = h * ((T2)BXF(0, 0, 0) + (T2)BXF(0, -1, 0));
    by = h * ((T2)BYF(0, 0, 0) + (T2)BYF(-1, 0, 0));
    bz = e * ((T2)BZF(0, 0, 0) + (T2)BZF(-1, 0, 0) + (T2)BZF(-1, -1, 0) + (T2)BZF(0, -1, 0) +
              (T2)BZF(0, 0, 1) + (T2)BZF(-1, 0, 1) + (T2)BZF(-1, -1, 1) + (T2)BZF(0, -1, 1));
  }
}

template <class T2, class Q2, bool FAST, class BXa, class BYa, class BZa>
__device__ __forceinline__ BBar<Q2> bbar(const BXa& BXF, const BYa& BYF, const BZa& BZF) {
  BBar<Q2> r;
  T2 x, y, z;
  bbar_point<0, T2>(BXF, BYF, BZF, x, y, z);
  r.x[0] = x; r.y[0] = y; r.z[0] = z;
  bbar_point<1, T2>(BXF, BYF, BZF, x, y, z);
  r.x[1] = x; r.y[1] = y; r.z[1] = z;
  bbar_point<2, T2>(BXF, BYF, BZF, x, y, z);
  r.x[2] = x; r.y[2] = y; r.z[2] = z;
#pragma unroll
  for (int p = 0; p < 3; ++p) r.m[p] = sq3<FAST>(r.x[p], r.y[p], r.z[p]);
  return r;
}

// E interpolated at E point P: its own component as is, the other two as
// 4-point averages (field_solver.F90:93-95, :116-118, :139-141; cleanEpar
// the same, :288-328).
template <int P, class T2, class AX, class AY, class AZ>
__device__ __forceinline__ void e_point(const AX& EX, const AY& EY, const AZ& EZ, T2& ex, T2& ey, T2& ez) {
  const T2 q = (T2)0.25;
  if constexpr (P == 0) {
    ex = (T2)EX(0, 0, 0);
    ey = q * ((T2)EY(0, 0, 0) + (T2)EY(1, 0, 0) + (T2)EY(0, -1, 0) + (T2)EY(1, -1, 0));
    ez = q * ((T2)EZ(0, 0, 0) + (T2)EZ(1, 0, 0) + (T2)EZ(0, 0, -1) + (T2)EZ(1, 0, -1));
  } else if constexpr (P == 1) {
    ex = q * ((T2)EX(0, 0, 0) + (T2)EX(-1, 0, 0) + (T2)EX(0, 1, 0) + (T2)EX(-1, 1, 0));
    ey = (T2)EY(0, 0, 0);
    ez = q * ((T2)EZ(0, 0, 0) + (T2)EZ(0, 1, 0) + (T2)EZ(0, 0, -1) + (T2)EZ(0, 1, -1));
  } else {
    ex = q * ((T2)EX(0, 0, 0) + (T2)EX(-1, 0, 0) + (T2)EX(0, 0, 1) + (T2)EX(-1, 0, 1));
    ey = q * ((T2)EY(0, 0, 0) + (T2)EY(0, -1, 0) + (T2)EY(0, 0, 1) + (T2)EY(0, -1, 1));
    ez = (T2)EZ(0, 0, 0);
  }
}

// 1/x in FP64 from the hardware estimate and two Newton steps (relative
// error ~1e-16; the near-null redo needs FP64 range and precision, not the
// correctly rounded quotient, and the full division sequence would set the
// register footprint of the whole FP32-bulk kernel)
__device__ __forceinline__ double rcp_nr(double x) {
  double r = __builtin_amdgcn_rcp(x);
  r = __builtin_fma(r, __builtin_fma(-x, r, 1.0), r);
  r = __builtin_fma(r, __builtin_fma(-x, r, 1.0), r);
  return r;
}

// cleanEpar of E point P in FP64 from the raw fields (the near-null redo,
// one point at a time to keep its register footprint small)
template <int P, class AX, class AY, class AZ, class BXa, class BYa, class BZa>
__device__ __forceinline__ double clean_point_fp64(const AX& EX, const AY& EY, const AZ& EZ, const BXa& BXF,
                                                   const BYa& BYF, const BZa& BZF, int guard) {
  double ex, ey, ez, bx, by, bz;
  e_point<P, double>(EX, EY, EZ, ex, ey, ez);
  bbar_point<P, double>(BXF, BYF, BZF, bx, by, bz);
  const double bmag = sq3<false>(bx, by, bz);
  const double epar = ex * bx + ey * by + ez * bz;
  const double e0 = P == 0 ? ex : (P == 1 ? ey : ez);
  const double b0 = P == 0 ? bx : (P == 1 ? by : bz);
  if (guard && bmag == 0.0) return e0;
  return e0 - epar * b0 * rcp_nr(bmag);
}

// The cleaned E component at E point P (field_solver.F90:288-328) given
// the B interpolants there (bx_, by_, bz_) and their |B|^2: E along its own
// axis at the point, the other two as 4-point averages in T2; epar = E.B in
// Q2, summed in the order x, y, z; E <- E - epar B / |B|^2 (no TINY).  With
// `guard`, |B|^2 == 0 leaves E unchanged.
template <int P, class T2, class Q2, bool FAST, class AX, class AY, class AZ>
__device__ __forceinline__ Q2 clean_point(const AX& EX, const AY& EY, const AZ& EZ, Q2 bx_, Q2 by_, Q2 bz_,
                                          Q2 bmag, int guard) {
  T2 intex, intey, intez;
  e_point<P, T2>(EX, EY, EZ, intex, intey, intez);
  const Q2 epar = FAST ? mad_<FAST>((Q2)intez, bz_, mad_<FAST>((Q2)intey, by_, (Q2)intex * bx_))
                       : ((Q2)intex * bx_ + (Q2)intey * by_ + (Q2)intez * bz_);
  const Q2 e0 = P == 0 ? (Q2)EX(0, 0, 0) : (P == 1 ? (Q2)EY(0, 0, 0) : (Q2)EZ(0, 0, 0));
  const Q2 b0 = P == 0 ? bx_ : (P == 1 ? by_ : bz_);
  return (guard && bmag == (Q2)0) ? e0 : (FAST ? e0 - div_<FAST>(epar * b0, bmag) : e0 - epar * b0 / bmag);
}

// cleanEpar at one cell (field_solver.F90:288-328) with sums in T2 and the
// projection in Q2, given the B interpolants: the three cleaned E
// components, clean_point at each E point.  `near` reports a point whose
// |B|^2 is below NEAR_NULL times `bref` (when bref > 0).
template <class T2, class Q2, bool FAST, class AX, class AY, class AZ>
__device__ __forceinline__ void clean_impl(const AX& EX, const AY& EY, const AZ& EZ, const BBar<Q2>& bb_in,
                                           int guard, Q2 bref, Q2& ox, Q2& oy, Q2& oz, bool& near) {
  near = false;
  // FAST: a point whose |B|^2 is non-zero but below TINY_B2 projects with
  // its B scaled by 2^64 (exactly; the projection does not depend on the
  // scale of B), so that the hardware reciprocal stays finite
  BBar<Q2> bb = bb_in;
  if constexpr (FAST) {
    // (an all-zero stencil, common where an absorbing layer has zeroed the
    // fields, stays out of the branch)
    if (fmin(fmin(bb.m[0], bb.m[1]), bb.m[2]) < (Q2)TINY_B2 && fmax(fmax(bb.m[0], bb.m[1]), bb.m[2]) > (Q2)0) {
#pragma unroll
      for (int p = 0; p < 3; ++p) {
        if (bb.m[p] > (Q2)0 && bb.m[p] < (Q2)TINY_B2) {
          const Q2 s = (Q2)0x1p64f;
          bb.x[p] *= s;
          bb.y[p] *= s;
          bb.z[p] *= s;
          bb.m[p] = sq3<FAST>(bb.x[p], bb.y[p], bb.z[p]);
        }
      }
    }
  }
#pragma unroll
  for (int p = 0; p < 3; ++p) near |= bb_in.m[p] < (Q2)NEAR_NULL * bref;
  ox = clean_point<0, T2, Q2, FAST>(EX, EY, EZ, bb.x[0], bb.y[0], bb.z[0], bb.m[0], guard);
  oy = clean_point<1, T2, Q2, FAST>(EX, EY, EZ, bb.x[1], bb.y[1], bb.z[1], bb.m[1], guard);
  oz = clean_point<2, T2, Q2, FAST>(EX, EY, EZ, bb.x[2], bb.y[2], bb.z[2], bb.m[2], guard);
}

// |B|^2 of the raw cell, the scale of the near-null test
template <class Q2, class BXa, class BYa, class BZa>
__device__ __forceinline__ Q2 cell_b2(const BXa& BXF, const BYa& BYF, const BZa& BZF) {
  const Q2 a = (Q2)BXF(0, 0, 0), b = (Q2)BYF(0, 0, 0), c = (Q2)BZF(0, 0, 0);
  return a * a + b * b + c * c;
}

// Whether the FP64 redo of clean_cell_bb or rhs_cell_bb applies at a cell:
// an interpolated |B|^2 below NEAR_NULL times the cell's own at one of the
// three E points (the test both use).
template <class Q>
__device__ __forceinline__ bool near_null_cell(const BBar<Q>& bb, Q bref) {
  bool r = false;
#pragma unroll
  for (int p = 0; p < 3; ++p) r |= bb.m[p] < (Q)NEAR_NULL * bref;
  return r;
}

// cleanEpar at one cell given its B interpolants `bb` and raw |B|^2 `bref`
// (0 unless PR::null_fp64); a near-null point redoes everything in FP64.
// REDO = false leaves out the redo (a caller that recomputes near-null cells
// elsewhere; near_null_cell names them).
template <class PR, bool REDO = PR::null_fp64, class AX, class AY, class AZ, class BXa, class BYa, class BZa>
__device__ __forceinline__ void clean_cell_bb(const AX& EX, const AY& EY, const AZ& EZ, const BXa& BXF,
                                              const BYa& BYF, const BZa& BZF, const BBar<typename PR::Q>& bb,
                                              typename PR::Q bref, int guard, typename PR::P& ox,
                                              typename PR::P& oy, typename PR::P& oz) {
  using T = typename PR::T;
  using Q = typename PR::Q;
  using P = typename PR::P;
  Q qx, qy, qz;
  bool near;
  clean_impl<T, Q, PR::null_fp64>(EX, EY, EZ, bb, guard, bref, qx, qy, qz, near);
  if (REDO && near) {
    ox = (P)clean_point_fp64<0>(EX, EY, EZ, BXF, BYF, BZF, guard);
    oy = (P)clean_point_fp64<1>(EX, EY, EZ, BXF, BYF, BZF, guard);
    oz = (P)clean_point_fp64<2>(EX, EY, EZ, BXF, BYF, BZF, guard);
    return;
  }
  ox = (P)qx;
  oy = (P)qy;
  oz = (P)qz;
}

template <class PR, class AX, class AY, class AZ, class BXa, class BYa, class BZa>
__device__ __forceinline__ void clean_cell(const AX& EX, const AY& EY, const AZ& EZ, const BXa& BXF,
                                           const BYa& BYF, const BZa& BZF, int guard,
                                           typename PR::P& ox, typename PR::P& oy,
                                           typename PR::P& oz) {
  using T = typename PR::T;
  using Q = typename PR::Q;
  const Q bref = PR::null_fp64 ? cell_b2<Q>(BXF, BYF, BZF) : (Q)0;
  const BBar<Q> bb = bbar<T, Q, PR::null_fp64>(BXF, BYF, BZF);
  clean_cell_bb<PR>(EX, EY, EZ, BXF, BYF, BZF, bb, bref, guard, ox, oy, oz);
}

// charge density at a node (computeRho, field_solver.F90:57-59)
template <class T, class AX, class AY, class AZ>
__device__ __forceinline__ T rho_node(const AX& EX, const AY& EY, const AZ& EZ, int a, int b, int d) {
  return (EX(a, b, d) - EX(a - 1, b, d)) + (EY(a, b, d) - EY(a, b - 1, d)) +
         (EZ(a, b, d) - EZ(a, b, d - 1));
}

// The three drift currents of rk3Push (field_solver.F90:90-157) with sums in
// T2 and the normalisation in Q2, given the B interpolants; j[] gets Jx, Jy,
// Jz.
template <class T2, class Q2, bool FAST, class AX, class AY, class AZ>
__device__ __forceinline__ void currents_impl(const AX& EX, const AY& EY, const AZ& EZ, const BBar<Q2>& bb,
                                              Q2 cc, T2 r000, T2 r100, T2 r010, T2 r001, Q2 bref, Q2* j,
                                              bool& near) {
  const T2 h = (T2)0.5;
  near = false;
  {  // X current (:90-111)
    T2 intrho = h * (r100 + r000);
    T2 intex, intey, intez;
    e_point<0, T2>(EX, EY, EZ, intex, intey, intez);
    j[0] = drift_current<Q2, FAST>(cc, (Q2)intrho, (Q2)intex, (Q2)intey, (Q2)intez, bb.x[0], bb.y[0],
                                   bb.z[0], bb.m[0], 0, bref, near);
  }
  {  // Y current (:113-134)
    T2 intrho = h * (r010 + r000);
    T2 intex, intey, intez;
    e_point<1, T2>(EX, EY, EZ, intex, intey, intez);
    j[1] = drift_current<Q2, FAST>(cc, (Q2)intrho, (Q2)intey, (Q2)intex, (Q2)intez, bb.y[1], bb.x[1],
                                   bb.z[1], bb.m[1], 1, bref, near);
  }
  {  // Z current (:136-157)
    T2 intrho = h * (r000 + r001);
    T2 intex, intey, intez;
    e_point<2, T2>(EX, EY, EZ, intex, intey, intez);
    j[2] = drift_current<Q2, FAST>(cc, (Q2)intrho, (Q2)intez, (Q2)intex, (Q2)intey, bb.z[2], bb.x[2],
                                   bb.y[2], bb.m[2], 2, bref, near);
  }
}

// Drift current at E point P in FP64 from the raw fields (the near-null
// redo of currents_impl, one component at a time)
template <int P, class AX, class AY, class AZ, class BXa, class BYa, class BZa>
__device__ __forceinline__ double current_point_fp64(const AX& EX, const AY& EY, const AZ& EZ, const BXa& BXF,
                                                     const BYa& BYF, const BZa& BZF, double cc,
                                                     double intrho) {
  double ex, ey, ez, bx, by, bz;
  e_point<P, double>(EX, EY, EZ, ex, ey, ez);
  bbar_point<P, double>(BXF, BYF, BZF, bx, by, bz);
  // the limiter and the drift current of drift_current, with the components
  // named by axis
  const double bmag = sq3<false>(bx, by, bz);
  const double emag = sq3<false>(ex, ey, ez);
  if (emag > bmag) {
    const double corr = sqrt(bmag / emag);  // rare branch: IEEE
    if (P != 0) ex *= corr;
    if (P != 1) ey *= corr;
    if (P != 2) ez *= corr;
  }
  double cross;
  if constexpr (P == 0)
    cross = ey * bz - by * ez;
  else if constexpr (P == 1)
    cross = ez * bx - ex * bz;
  else
    cross = ex * by - bx * ey;
  return cc * (intrho * cross) * rcp_nr(bmag + 1e-7);  // TINY, src/defs.F90:4
}

// The drift currents of rk3Push at one cell given the B interpolants `bb`
// and raw |B|^2 `bref` of the cell.  RHO(a, b, d) gives the charge density
// at node (i+a, j+b, k+d).  REDO: a near-null cell redoes them in FP64 from
// the raw B (BXF, BYF, BZF are read only then).
template <class PR, bool REDO, class AX, class AY, class AZ, class BXa, class BYa, class BZa, class RF>
__device__ __forceinline__ void currents_of(const AX& EX, const AY& EY, const AZ& EZ, const BXa& BXF,
                                            const BYa& BYF, const BZa& BZF, const RF& RHO,
                                            const BBar<typename PR::Q>& bb, typename PR::Q bref,
                                            typename PR::P cc, typename PR::P& jx, typename PR::P& jy,
                                            typename PR::P& jz) {
  using T = typename PR::T;
  using Q = typename PR::Q;
  using P = typename PR::P;
  const T r000 = RHO(0, 0, 0);
  const T r100 = RHO(1, 0, 0);
  const T r010 = RHO(0, 1, 0);
  const T r001 = RHO(0, 0, 1);
  Q j[3];
  bool near;
  currents_impl<T, Q, PR::null_fp64>(EX, EY, EZ, bb, (Q)cc, r000, r100, r010, r001, bref, j, near);
  if constexpr (REDO) {
    if (near) {
      // FP64 redo, one current at a time (small register footprint)
      const double c = (double)cc, h = 0.5;
      const double q000 = (double)r000;
      jx = (P)current_point_fp64<0>(EX, EY, EZ, BXF, BYF, BZF, c, h * ((double)r100 + q000));
      jy = (P)current_point_fp64<1>(EX, EY, EZ, BXF, BYF, BZF, c, h * ((double)r010 + q000));
      jz = (P)current_point_fp64<2>(EX, EY, EZ, BXF, BYF, BZF, c, h * (q000 + (double)r001));
      return;
    }
  }
  jx = (P)j[0];
  jy = (P)j[1];
  jz = (P)j[2];
}

// The raw-B curl brackets of rk3Push at a cell (field_solver.F90:84-89):
// cb[0] = (by(k) - by(k-1)) - (bz(j) - bz(j-1)), and cyclically.
template <class T, class BXa, class BYa, class BZa>
__device__ __forceinline__ void curl_b(const BXa& BXF, const BYa& BYF, const BZa& BZF, T* cb) {
  cb[0] = (BYF(0, 0, 0) - BYF(0, 0, -1)) - (BZF(0, 0, 0) - BZF(0, -1, 0));
  cb[1] = (BZF(0, 0, 0) - BZF(-1, 0, 0)) - (BXF(0, 0, 0) - BXF(0, 0, -1));
  cb[2] = (BXF(0, 0, 0) - BXF(0, -1, 0)) - (BYF(0, 0, 0) - BYF(-1, 0, 0));
}

// dE and dB of rk3Push (field_solver.F90:81-89, :160-162) from the
// currents, the raw-B curl brackets cb and the B0 curl bracket t:
// d[0..2] = dE, d[3..5] = dB.
template <class PR, class AX, class AY, class AZ>
__device__ __forceinline__ void rhs_tail(const AX& EX, const AY& EY, const AZ& EZ, typename PR::P cc,
                                         typename PR::T tx, typename PR::T ty, typename PR::T tz,
                                         const typename PR::T* cb, typename PR::P jx, typename PR::P jy,
                                         typename PR::P jz, typename PR::P* d) {
  using P = typename PR::P;
  const P dbx = cc * (P)((EY(0, 0, 1) - EY(0, 0, 0)) - (EZ(0, 1, 0) - EZ(0, 0, 0)));
  const P dby = cc * (P)((EZ(1, 0, 0) - EZ(0, 0, 0)) - (EX(0, 0, 1) - EX(0, 0, 0)));
  const P dbz = cc * (P)((EX(0, 1, 0) - EX(0, 0, 0)) - (EY(1, 0, 0) - EY(0, 0, 0)));
  const P dex = -cc * (P)(cb[0] - tx) - jx;
  const P dey = -cc * (P)(cb[1] - ty) - jy;
  const P dez = -cc * (P)(cb[2] - tz) - jz;
  d[0] = dex; d[1] = dey; d[2] = dez;
  d[3] = dbx; d[4] = dby; d[5] = dbz;
}

// rk3Push at one cell (field_solver.F90:81-162) given the B interpolants
// `bb` and raw |B|^2 `bref` of the cell: d[0..2] = dE, d[3..5] = dB.
// RHO(a, b, d) gives the charge density at node (i+a, j+b, k+d); t0 is the
// precomputed B0 curl bracket (zero for a uniform background).
template <class PR, bool REDO = PR::null_fp64, class AX, class AY, class AZ, class BXa, class BYa, class BZa,
          class RF>
__device__ __forceinline__ void rhs_cell_bb(const AX& EX, const AY& EY, const AZ& EZ, const BXa& BXF,
                                            const BYa& BYF, const BZa& BZF, const RF& RHO,
                                            const BBar<typename PR::Q>& bb, typename PR::Q bref,
                                            typename PR::P cc, typename PR::T tx, typename PR::T ty,
                                            typename PR::T tz, typename PR::P* d) {
  using T = typename PR::T;
  using P = typename PR::P;
  // the currents first: their FP64 redo then runs with the fewest other
  // values live
  P jx, jy, jz;
  currents_of<PR, REDO>(EX, EY, EZ, BXF, BYF, BZF, RHO, bb, bref, cc, jx, jy, jz);
  T cb[3];
  curl_b<T>(BXF, BYF, BZF, cb);
  rhs_tail<PR>(EX, EY, EZ, cc, tx, ty, tz, cb, jx, jy, jz, d);
}

// rk3Push at one cell from precomputed raw-B terms: the B interpolants, the
// raw |B|^2 and the curl brackets of the cell's raw B (no near-null redo).
template <class PR, class AX, class AY, class AZ, class RF>
__device__ __forceinline__ void rhs_cell_pre(const AX& EX, const AY& EY, const AZ& EZ, const RF& RHO,
                                             const BBar<typename PR::Q>& bb, typename PR::Q bref,
                                             typename PR::P cc, typename PR::T tx, typename PR::T ty,
                                             typename PR::T tz, const typename PR::T* cb, typename PR::P* d) {
  using T = typename PR::T;
  using P = typename PR::P;
  auto none = [](int, int, int) { return (T)0; };
  P jx, jy, jz;
  currents_of<PR, false>(EX, EY, EZ, none, none, none, RHO, bb, bref, cc, jx, jy, jz);
  rhs_tail<PR>(EX, EY, EZ, cc, tx, ty, tz, cb, jx, jy, jz, d);
}

template <class PR, class AX, class AY, class AZ, class BXa, class BYa, class BZa, class RF>
__device__ __forceinline__ void rhs_cell(const AX& EX, const AY& EY, const AZ& EZ, const BXa& BXF,
                                         const BYa& BYF, const BZa& BZF, const RF& RHO,
                                         typename PR::P cc, typename PR::T tx, typename PR::T ty,
                                         typename PR::T tz, typename PR::P* d) {
  using T = typename PR::T;
  using Q = typename PR::Q;
  const Q bref = PR::null_fp64 ? cell_b2<Q>(BXF, BYF, BZF) : (Q)0;
  const BBar<Q> bb = bbar<T, Q, PR::null_fp64>(BXF, BYF, BZF);
  rhs_cell_bb<PR>(EX, EY, EZ, BXF, BYF, BZF, RHO, bb, bref, cc, tx, ty, tz, d);
}

// rk3Update (field_solver.F90:187-193 and the absorbing layer :200-265) at
// one interior cell: v holds X on entry (the c2 term), u holds U^n; on exit
// v holds the new values in working precision, ready to round to S.
template <class PR>
__device__ __forceinline__ void update_cell(typename PR::P* v, const typename PR::P* u,
                                            const typename PR::P* d, typename PR::P c1,
                                            typename PR::P c2, typename PR::P c3,
                                            const Absorb<typename PR::P>& ab, int i, int j, int k) {
  using S = typename PR::S;
  using P = typename PR::P;
  if (PR::null_fp64 && c2 == c3 && c1 + c2 == (P)1) {
    // the SSP-RK3 stages all have c1 = 1 - c and c2 = c3 = c: U^n + c ((X -
    // U^n) + dU) keeps the state's scale exact, where rounded 1/3 and 2/3
    // would drift it by their rounding every step
    for (int m = 0; m < 6; ++m) v[m] = mad_<true>(c2, (v[m] - u[m]) + d[m], u[m]);
  } else {
    for (int m = 0; m < 6; ++m)
      v[m] = PR::null_fp64 ? mad_<true>(c3, d[m], mad_<true>(c2, v[m], c1 * u[m]))
                           : c1 * u[m] + c2 * v[m] + c3 * d[m];
  }
  constexpr bool fast_layer = PR::null_fp64 && sizeof(P) == 4;
  if (fast_layer && ab.on) {
    // FP32-bulk: the same layer from six squared distances that share
    // their per-axis squares, hardware square roots (1 ulp) and a multiply
    // by 1/buff (S = P here, so there is no storage rounding to apply)
    const P ig = (P)i, jg = ab.fixed ? (P)j : (P)i, kg = (P)k;
    const P x0 = ig - ab.xc, y0 = jg - ab.yc, z0 = kg - ab.zc;
    const P x1 = x0 + (P)0.5, y1 = y0 + (P)0.5, z1 = z0 + (P)0.5;
    const P xx0 = x0 * x0, yy0 = y0 * y0, zz0 = z0 * z0, xx1 = x1 * x1, yy1 = y1 * y1, zz1 = z1 * z1;
    const P lim = ab.rabsorb - (P)1.0;
    if (lim < (P)0.0 || xx0 + yy0 + zz0 > lim * lim) {
      const P ib = __builtin_amdgcn_rcpf(ab.buff);
      // Yee offsets of ex ey ez bx by bz
      const P r2[6] = {xx1 + yy0 + zz0, xx0 + yy1 + zz0, xx0 + yy0 + zz1,
                       xx0 + yy1 + zz1, xx1 + yy0 + zz1, xx1 + yy1 + zz0};
#pragma unroll
      for (int m = 0; m < 6; ++m) {
        P dd = (__builtin_amdgcn_sqrtf(r2[m]) - ab.rabsorb) * ib;
        if (ab.fixed) dd = dd < (P)1.0 ? dd : (P)1.0;
        if (dd > (P)0.0) v[m] = mad_<true>(-dd * dd, dd * v[m], v[m]);
      }
    }
  } else if (!fast_layer && ab.on) {
    // the reference stores first and damps the stored value
    for (int m = 0; m < 6; ++m) v[m] = (P)(S)v[m];
    const P ig = (P)i, jg = ab.fixed ? (P)j : (P)i, kg = (P)k;
    P r = sqrt((ig - ab.xc) * (ig - ab.xc) + (jg - ab.yc) * (jg - ab.yc) + (kg - ab.zc) * (kg - ab.zc));
    if (r > ab.rabsorb - (P)1.0) {
      const P hh = (P)0.5, zz = (P)0.0;
      // Yee offsets of ex ey ez bx by bz
      const P ox[6] = {hh, zz, zz, zz, hh, hh};
      const P oy[6] = {zz, hh, zz, hh, zz, hh};
      const P oz[6] = {zz, zz, hh, hh, hh, zz};
#pragma unroll
      for (int m = 0; m < 6; ++m) {
        r = sqrt((ig + ox[m] - ab.xc) * (ig + ox[m] - ab.xc) + (jg + oy[m] - ab.yc) * (jg + oy[m] - ab.yc) +
                 (kg + oz[m] - ab.zc) * (kg + oz[m] - ab.zc));
        P dd = (r - ab.rabsorb) / ab.buff;
        if (ab.fixed && dd > (P)1.0) dd = (P)1.0;
        if (dd > (P)0.0) v[m] = ((P)1.0 - dd * dd * dd) * v[m];
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Compensated state (TFF_FP32_COMP): every field value is an unevaluated
// FP32 pair hi + lo with |lo| <= ulp(hi) / 2; stencils read hi, the stage
// accumulation and the projection carry lo.  The pair arithmetic rests on
// Knuth's two-sum, an error-free transformation in round-to-nearest (the
// library is built with -ffp-contract=off, so the compiler keeps its four
// roundings apart).

// s + e = a + b exactly, s = fl(a + b)
__device__ __forceinline__ void two_sum(float a, float b, float& s, float& e) {
  s = a + b;
  const float bv = s - a;
  e = (a - (s - bv)) + (b - bv);
}

// The SSP-RK3 stage Y = U + c2 (X - U) + c3 dU (c1 = 1 - c2) on pairs:
// X_hi - U_hi exactly by two-sum, the increment in FP32 (its own rounding
// is relative to the increment, not to U), and one two-sum onto U_hi.
__device__ __forceinline__ void comp_update(float uh, float ul, float xh, float xl, float d, float c2, float c3,
                                            float& yh, float& yl) {
  float dh, de;
  two_sum(xh, -uh, dh, de);
  const float dl = de + (xl - ul);
  const float inc = __builtin_fmaf(c2, dh, __builtin_fmaf(c2, dl, c3 * d));
  two_sum(uh, ul + inc, yh, yl);
}

// (yh + yl) * f on pairs: the product's rounding error by fma (two-product)
__device__ __forceinline__ void comp_scale(float& yh, float& yl, float f) {
  const float p = yh * f;
  const float e = __builtin_fmaf(yh, f, -p);
  two_sum(p, __builtin_fmaf(yl, f, e), yh, yl);
}

// Damping factor of the absorbing layer for component m (Yee offsets of
// ex ey ez bx by bz) at cell (i, j, k), 1 outside the layer; FP32 with
// hardware square roots, as update_cell's FP32-bulk layer.
__device__ __forceinline__ float absorb_factor(const Absorb<float>& ab, int i, int j, int k, int m) {
  const float hx[6] = {0.5f, 0.f, 0.f, 0.f, 0.5f, 0.5f};
  const float hy[6] = {0.f, 0.5f, 0.f, 0.5f, 0.f, 0.5f};
  const float hz[6] = {0.f, 0.f, 0.5f, 0.5f, 0.5f, 0.f};
  const float ig = (float)i, jg = ab.fixed ? (float)j : (float)i, kg = (float)k;
  const float x0 = ig - ab.xc, y0 = jg - ab.yc, z0 = kg - ab.zc;
  const float lim = ab.rabsorb - 1.0f;
  if (!(lim < 0.0f || x0 * x0 + y0 * y0 + z0 * z0 > lim * lim)) return 1.0f;
  const float x = x0 + hx[m], y = y0 + hy[m], z = z0 + hz[m];
  float dd = (__builtin_amdgcn_sqrtf(x * x + y * y + z * z) - ab.rabsorb) * __builtin_amdgcn_rcpf(ab.buff);
  if (ab.fixed) dd = dd < 1.0f ? dd : 1.0f;
  return dd > 0.0f ? 1.0f - dd * dd * dd : 1.0f;
}

}  // namespace tff
