// tff_output.h -- node-interpolated fields for the output path.
//
// writeFields_hdf5 (src/io/write_output.F90:185-429) stores, at every
// output_istep-th node of the meshblock, E interpolated from the cell edges
// and B from the faces (interpFromEdges / interpFromFaces at offset 0,
// src/tools/helpers.F90:93-257).  k_output does the sampling and the
// interpolation on the device, in the storage precision, with the
// reference's operation order; the host writes the arrays out.
#pragma once
#include "tff_kernels.h"

namespace tff {

// (di, dj, dk) of the terms of each corner average, corners in the order
// c000 c100 c010 c110 c001 c101 c011 c111 and terms in the reference's
// order (a corner's sum rounds in that order): E components from the
// edges (helpers.F90:113-154), B components from the faces (:184-241).
struct Off3 {
  signed char i, j, k;
};
__constant__ constexpr Off3 kEdgeTerms[3][8][2] = {
    {{{0, 0, 0}, {-1, 0, 0}}, {{0, 0, 0}, {1, 0, 0}}, {{0, 1, 0}, {-1, 1, 0}}, {{0, 1, 0}, {1, 1, 0}},
     {{0, 0, 1}, {-1, 0, 1}}, {{0, 0, 1}, {1, 0, 1}}, {{0, 1, 1}, {-1, 1, 1}}, {{0, 1, 1}, {1, 1, 1}}},
    {{{0, 0, 0}, {0, -1, 0}}, {{1, 0, 0}, {1, -1, 0}}, {{0, 0, 0}, {0, 1, 0}}, {{1, 0, 0}, {1, 1, 0}},
     {{0, 0, 1}, {0, -1, 1}}, {{1, 0, 1}, {1, -1, 1}}, {{0, 0, 1}, {0, 1, 1}}, {{1, 0, 1}, {1, 1, 1}}},
    {{{0, 0, 0}, {0, 0, -1}}, {{1, 0, 0}, {1, 0, -1}}, {{0, 1, 0}, {0, 1, -1}}, {{1, 1, 0}, {1, 1, -1}},
     {{0, 0, 0}, {0, 0, 1}}, {{1, 0, 0}, {1, 0, 1}}, {{0, 1, 0}, {0, 1, 1}}, {{1, 1, 0}, {1, 1, 1}}},
};
__constant__ constexpr Off3 kFaceTerms[3][8][4] = {
    {{{0, 0, 0}, {0, -1, 0}, {0, 0, -1}, {0, -1, -1}},
     {{1, 0, 0}, {1, -1, 0}, {1, 0, -1}, {1, -1, -1}},
     {{0, 0, 0}, {0, 1, 0}, {0, 0, -1}, {0, 1, -1}},
     {{1, 0, 0}, {1, 0, -1}, {1, 1, -1}, {1, 1, 0}},
     {{0, 0, 0}, {0, 0, 1}, {0, -1, 0}, {0, -1, 1}},
     {{1, 0, 0}, {1, 0, 1}, {1, -1, 0}, {1, -1, 1}},
     {{0, 0, 0}, {0, 1, 0}, {0, 1, 1}, {0, 0, 1}},
     {{1, 0, 0}, {1, 1, 0}, {1, 1, 1}, {1, 0, 1}}},
    {{{-1, 0, -1}, {-1, 0, 0}, {0, 0, -1}, {0, 0, 0}},
     {{0, 0, -1}, {0, 0, 0}, {1, 0, -1}, {1, 0, 0}},
     {{-1, 1, -1}, {-1, 1, 0}, {0, 1, -1}, {0, 1, 0}},
     {{0, 1, -1}, {0, 1, 0}, {1, 1, -1}, {1, 1, 0}},
     {{-1, 0, 0}, {-1, 0, 1}, {0, 0, 0}, {0, 0, 1}},
     {{0, 0, 0}, {0, 0, 1}, {1, 0, 0}, {1, 0, 1}},
     {{-1, 1, 0}, {-1, 1, 1}, {0, 1, 0}, {0, 1, 1}},
     {{0, 1, 0}, {0, 1, 1}, {1, 1, 0}, {1, 1, 1}}},
    {{{-1, -1, 0}, {-1, 0, 0}, {0, -1, 0}, {0, 0, 0}},
     {{0, -1, 0}, {0, 0, 0}, {1, -1, 0}, {1, 0, 0}},
     {{-1, 0, 0}, {-1, 1, 0}, {0, 0, 0}, {0, 1, 0}},
     {{0, 0, 0}, {0, 1, 0}, {1, 0, 0}, {1, 1, 0}},
     {{-1, -1, 1}, {-1, 0, 1}, {0, -1, 1}, {0, 0, 1}},
     {{0, -1, 1}, {0, 0, 1}, {1, -1, 1}, {1, 0, 1}},
     {{-1, 0, 1}, {-1, 1, 1}, {0, 0, 1}, {0, 1, 1}},
     {{0, 0, 1}, {0, 1, 1}, {1, 0, 1}, {1, 1, 1}}},
};

// The trilinear blend of the eight corner values at offset (dx, dy, dz):
// along x (c00 c10 c01 c11), then y (c0 c1), then z.  At offset 0 it still
// forms every product, as the reference does (a -0 corner or a non-finite
// neighbour shows in the result the same way).
template <class T>
__device__ __forceinline__ T trilinear(const T* c, T dx, T dy, T dz) {
  const T c00 = c[0] * ((T)1 - dx) + c[1] * dx;
  const T c10 = c[2] * ((T)1 - dx) + c[3] * dx;
  const T c01 = c[4] * ((T)1 - dx) + c[5] * dx;
  const T c11 = c[6] * ((T)1 - dx) + c[7] * dx;
  const T c0 = c00 * ((T)1 - dy) + c10 * dy;
  const T c1 = c01 * ((T)1 - dy) + c11 * dy;
  return c0 * ((T)1 - dz) + c1 * dz;
}

// interpFromEdges (a = 0..2 -> E component a) and interpFromFaces (B) at
// offset (dx, dy, dz) from the cell at element c of a component array
// with row pitch sy and plane pitch sz.
template <class S>
__device__ __forceinline__ S interp_edges(const S* f, int64_t c, int64_t sy, int64_t sz, int a, S dx, S dy,
                                          S dz) {
  S v[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const Off3 p = kEdgeTerms[a][q][0], r = kEdgeTerms[a][q][1];
    v[q] = (S)0.5 * (f[c + p.i + p.j * sy + p.k * sz] + f[c + r.i + r.j * sy + r.k * sz]);
  }
  return trilinear(v, dx, dy, dz);
}
template <class S>
__device__ __forceinline__ S interp_faces(const S* f, int64_t c, int64_t sy, int64_t sz, int a, S dx, S dy,
                                          S dz) {
  S v[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    S s = (S)0;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const Off3 p = kFaceTerms[a][q][t];
      const S x = f[c + p.i + p.j * sy + p.k * sz];
      s = t == 0 ? x : s + x;
    }
    v[q] = (S)0.25 * s;
  }
  return trilinear(v, dx, dy, dz);
}

// One thread per output node (i1, j1, k1) at cell (i1, j1, k1) * istep:
// out[f * nout + i1 + ni * (j1 + nj * k1)] for f = ex ey ez bx by bz, in
// the host type O.
template <class S, class O>
__global__ void __launch_bounds__(256) k_output(Geom g, int istep, int ni, int nj, int nk, V3<const S> e,
                                                V3<const S> b, O* out) {
  const int i1 = blockIdx.x * 64 + threadIdx.x % 64;
  const int j1 = blockIdx.y * 4 + threadIdx.x / 64;
  const int k1 = blockIdx.z;
  if (i1 >= ni || j1 >= nj || k1 >= nk) return;
  const int64_t c = g.at(i1 * istep, j1 * istep, k1 * istep);
  const int64_t sy = g.px, sz = g.plane;
  const int64_t nout = (int64_t)ni * nj * nk;
  const int64_t t = i1 + (int64_t)ni * (j1 + (int64_t)nj * k1);
  const S zero = (S)0;
  const S* ef[3] = {e.x, e.y, e.z};
  const S* bf[3] = {b.x, b.y, b.z};
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    out[a * nout + t] = (O)interp_edges(ef[a], c, sy, sz, a, zero, zero, zero);
    out[(3 + a) * nout + t] = (O)interp_faces(bf[a], c, sy, sz, a, zero, zero, zero);
  }
}

// The values the user files' userOutput reads at the output nodes
// (write_output.F90:375-383, e.g. user_uniform.F90:151-173): the raw
// components ex..bz at cell (i1, j1, k1) * istep, and the charge density
// there (computeRho's difference form, field_solver.F90:57-59, of the
// state being written) as the seventh dataset.
template <class S, class O>
__global__ void __launch_bounds__(256) k_output_samples(Geom g, int istep, int ni, int nj, int nk,
                                                        V3<const S> e, V3<const S> b, O* out) {
  const int i1 = blockIdx.x * 64 + threadIdx.x % 64;
  const int j1 = blockIdx.y * 4 + threadIdx.x / 64;
  const int k1 = blockIdx.z;
  if (i1 >= ni || j1 >= nj || k1 >= nk) return;
  const int64_t c = g.at(i1 * istep, j1 * istep, k1 * istep);
  const int64_t sy = g.px, sz = g.plane;
  const int64_t nout = (int64_t)ni * nj * nk;
  const int64_t t = i1 + (int64_t)ni * (j1 + (int64_t)nj * k1);
  out[t] = (O)e.x[c];
  out[nout + t] = (O)e.y[c];
  out[2 * nout + t] = (O)e.z[c];
  out[3 * nout + t] = (O)b.x[c];
  out[4 * nout + t] = (O)b.y[c];
  out[5 * nout + t] = (O)b.z[c];
  out[6 * nout + t] = (O)((e.x[c] - e.x[c - 1]) + (e.y[c] - e.y[c - sy]) + (e.z[c] - e.z[c - sz]));
}

}  // namespace tff
