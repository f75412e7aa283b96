// tff_cell.h -- per-cell arithmetic of the force-free step, written once and
// shared by every kernel (global-memory and LDS-tiled forms), so that all
// paths evaluate the same expression trees in the same order.
//
// Field access goes through accessors f(a, b, d) returning the value at
// offset (a, b, d) from the cell, already converted to the stencil type T.
#pragma once
#include "tff_common.h"

namespace tff {

// cleanEpar at one cell (field_solver.F90:288-328): returns the three
// cleaned E components.  Interpolants: E along its own axis at the point,
// the other two as 4-point averages, B as one 8-point and two 2-point
// averages; epar = E.B; E <- E - epar B / |B|^2 (no TINY).  With `guard`,
// |B|^2 == 0 leaves E unchanged.
template <class PR, class AX, class AY, class AZ, class BXa, class BYa, class BZa>
__device__ __forceinline__ void clean_cell(const AX& EX, const AY& EY, const AZ& EZ, const BXa& BXF,
                                           const BYa& BYF, const BZa& BZF, int guard,
                                           typename PR::P& ox, typename PR::P& oy,
                                           typename PR::P& oz) {
  using T = typename PR::T;
  using P = typename PR::P;
  const T h = (T)0.5, q = (T)0.25, e = (T)0.125;
  {
    T intex = EX(0, 0, 0);
    T intey = q * (EY(0, 0, 0) + EY(1, 0, 0) + EY(0, -1, 0) + EY(1, -1, 0));
    T intez = q * (EZ(0, 0, 0) + EZ(1, 0, 0) + EZ(0, 0, -1) + EZ(1, 0, -1));
    T intbx = e * (BXF(0, 0, 0) + BXF(0, -1, 0) + BXF(1, -1, 0) + BXF(1, 0, 0) + BXF(0, 0, -1) +
                   BXF(0, -1, -1) + BXF(1, -1, -1) + BXF(1, 0, -1));
    T intby = h * (BYF(0, 0, 0) + BYF(0, 0, -1));
    T intbz = h * (BZF(0, 0, 0) + BZF(0, -1, 0));
    P bx_ = intbx, by_ = intby, bz_ = intbz;
    P bmag = bx_ * bx_ + by_ * by_ + bz_ * bz_;
    P epar = ((P)intex * bx_ + (P)intey * by_ + (P)intez * bz_);
    const P e0 = (P)EX(0, 0, 0);
    ox = (guard && bmag == (P)0) ? e0 : e0 - epar * bx_ / bmag;
  }
  {
    T intex = q * (EX(0, 0, 0) + EX(-1, 0, 0) + EX(0, 1, 0) + EX(-1, 1, 0));
    T intey = EY(0, 0, 0);
    T intez = q * (EZ(0, 0, 0) + EZ(0, 1, 0) + EZ(0, 0, -1) + EZ(0, 1, -1));
    T intbx = h * (BXF(0, 0, 0) + BXF(0, 0, -1));
    T intby = e * (BYF(0, 0, 0) + BYF(-1, 0, 0) + BYF(-1, 1, 0) + BYF(0, 1, 0) + BYF(0, 0, -1) +
                   BYF(-1, 0, -1) + BYF(-1, 1, -1) + BYF(0, 1, -1));
    T intbz = h * (BZF(0, 0, 0) + BZF(-1, 0, 0));
    P bx_ = intbx, by_ = intby, bz_ = intbz;
    P bmag = bx_ * bx_ + by_ * by_ + bz_ * bz_;
    P epar = ((P)intex * bx_ + (P)intey * by_ + (P)intez * bz_);
    const P e0 = (P)EY(0, 0, 0);
    oy = (guard && bmag == (P)0) ? e0 : e0 - epar * by_ / bmag;
  }
  {
    T intex = q * (EX(0, 0, 0) + EX(-1, 0, 0) + EX(0, 0, 1) + EX(-1, 0, 1));
    T intey = q * (EY(0, 0, 0) + EY(0, -1, 0) + EY(0, 0, 1) + EY(0, -1, 1));
    T intez = EZ(0, 0, 0);
    T intbx = h * (BXF(0, 0, 0) + BXF(0, -1, 0));
    T intby = h * (BYF(0, 0, 0) + BYF(-1, 0, 0));
    T intbz = e * (BZF(0, 0, 0) + BZF(-1, 0, 0) + BZF(-1, -1, 0) + BZF(0, -1, 0) + BZF(0, 0, 1) +
                   BZF(-1, 0, 1) + BZF(-1, -1, 1) + BZF(0, -1, 1));
    P bx_ = intbx, by_ = intby, bz_ = intbz;
    P bmag = bx_ * bx_ + by_ * by_ + bz_ * bz_;
    P epar = ((P)intex * bx_ + (P)intey * by_ + (P)intez * bz_);
    const P e0 = (P)EZ(0, 0, 0);
    oz = (guard && bmag == (P)0) ? e0 : e0 - epar * bz_ / bmag;
  }
}

// charge density at a node (computeRho, field_solver.F90:57-59)
template <class T, class AX, class AY, class AZ>
__device__ __forceinline__ T rho_node(const AX& EX, const AY& EY, const AZ& EZ, int a, int b, int d) {
  return (EX(a, b, d) - EX(a - 1, b, d)) + (EY(a, b, d) - EY(a, b - 1, d)) +
         (EZ(a, b, d) - EZ(a, b, d - 1));
}

// rk3Push at one cell (field_solver.F90:81-162): d[0..2] = dE, d[3..5] = dB.
// RHO(a, b, d) gives the charge density at node (i+a, j+b, k+d); t0 is the
// precomputed B0 curl bracket (zero for a uniform background).
template <class PR, class AX, class AY, class AZ, class BXa, class BYa, class BZa, class RF>
__device__ __forceinline__ void rhs_cell(const AX& EX, const AY& EY, const AZ& EZ, const BXa& BXF,
                                         const BYa& BYF, const BZa& BZF, const RF& RHO,
                                         typename PR::P cc, typename PR::T tx, typename PR::T ty,
                                         typename PR::T tz, typename PR::P* d) {
  using T = typename PR::T;
  using P = typename PR::P;
  P dbx = cc * (P)((EY(0, 0, 1) - EY(0, 0, 0)) - (EZ(0, 1, 0) - EZ(0, 0, 0)));
  P dby = cc * (P)((EZ(1, 0, 0) - EZ(0, 0, 0)) - (EX(0, 0, 1) - EX(0, 0, 0)));
  P dbz = cc * (P)((EX(0, 1, 0) - EX(0, 0, 0)) - (EY(1, 0, 0) - EY(0, 0, 0)));
  P dex = -cc * (P)(((BYF(0, 0, 0) - BYF(0, 0, -1)) - (BZF(0, 0, 0) - BZF(0, -1, 0))) - tx);
  P dey = -cc * (P)(((BZF(0, 0, 0) - BZF(-1, 0, 0)) - (BXF(0, 0, 0) - BXF(0, 0, -1))) - ty);
  P dez = -cc * (P)(((BXF(0, 0, 0) - BXF(0, -1, 0)) - (BYF(0, 0, 0) - BYF(-1, 0, 0))) - tz);
  const T r000 = RHO(0, 0, 0);
  const T r100 = RHO(1, 0, 0);
  const T r010 = RHO(0, 1, 0);
  const T r001 = RHO(0, 0, 1);
  const T h = (T)0.5, q = (T)0.25, e = (T)0.125;
  {  // X current (:90-111)
    T intrho = h * (r100 + r000);
    T intex = EX(0, 0, 0);
    T intey = q * (EY(0, 0, 0) + EY(1, 0, 0) + EY(0, -1, 0) + EY(1, -1, 0));
    T intez = q * (EZ(0, 0, 0) + EZ(1, 0, 0) + EZ(0, 0, -1) + EZ(1, 0, -1));
    T intbx = e * (BXF(0, 0, 0) + BXF(0, -1, 0) + BXF(1, -1, 0) + BXF(1, 0, 0) + BXF(0, 0, -1) +
                   BXF(0, -1, -1) + BXF(1, -1, -1) + BXF(1, 0, -1));
    T intby = h * (BYF(0, 0, 0) + BYF(0, 0, -1));
    T intbz = h * (BZF(0, 0, 0) + BZF(0, -1, 0));
    dex = dex - drift_current<P>(cc, (P)intrho, (P)intex, (P)intey, (P)intez, (P)intbx, (P)intby,
                                 (P)intbz, 0);
  }
  {  // Y current (:113-134)
    T intrho = h * (r010 + r000);
    T intex = q * (EX(0, 0, 0) + EX(-1, 0, 0) + EX(0, 1, 0) + EX(-1, 1, 0));
    T intey = EY(0, 0, 0);
    T intez = q * (EZ(0, 0, 0) + EZ(0, 1, 0) + EZ(0, 0, -1) + EZ(0, 1, -1));
    T intbx = h * (BXF(0, 0, 0) + BXF(0, 0, -1));
    T intby = e * (BYF(0, 0, 0) + BYF(-1, 0, 0) + BYF(-1, 1, 0) + BYF(0, 1, 0) + BYF(0, 0, -1) +
                   BYF(-1, 0, -1) + BYF(-1, 1, -1) + BYF(0, 1, -1));
    T intbz = h * (BZF(0, 0, 0) + BZF(-1, 0, 0));
    dey = dey - drift_current<P>(cc, (P)intrho, (P)intey, (P)intex, (P)intez, (P)intby, (P)intbx,
                                 (P)intbz, 1);
  }
  {  // Z current (:136-157)
    T intrho = h * (r000 + r001);
    T intex = q * (EX(0, 0, 0) + EX(-1, 0, 0) + EX(0, 0, 1) + EX(-1, 0, 1));
    T intey = q * (EY(0, 0, 0) + EY(0, -1, 0) + EY(0, 0, 1) + EY(0, -1, 1));
    T intez = EZ(0, 0, 0);
    T intbx = h * (BXF(0, 0, 0) + BXF(0, -1, 0));
    T intby = h * (BYF(0, 0, 0) + BYF(-1, 0, 0));
    T intbz = e * (BZF(0, 0, 0) + BZF(-1, 0, 0) + BZF(-1, -1, 0) + BZF(0, -1, 0) + BZF(0, 0, 1) +
                   BZF(-1, 0, 1) + BZF(-1, -1, 1) + BZF(0, -1, 1));
    dez = dez - drift_current<P>(cc, (P)intrho, (P)intez, (P)intex, (P)intey, (P)intbz, (P)intbx,
                                 (P)intby, 2);
  }
  d[0] = dex; d[1] = dey; d[2] = dez;
  d[3] = dbx; d[4] = dby; d[5] = dbz;
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
  for (int m = 0; m < 6; ++m) v[m] = c1 * u[m] + c2 * v[m] + c3 * d[m];
  if (ab.on) {
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

}  // namespace tff
