// tff_fused.h -- the fused stage kernel of the fast path.
//
// One launch per RK stage.  It reads the raw output of the previous stage
// (after its boundary hook and ghost fill), applies that stage's cleanEpar
// on the fly, evaluates rk3Push + rk3Update on the cleaned fields and
// writes the raw output of this stage.  The clean therefore costs no HBM
// pass of its own: per cell-update the three launches move 15 + 18 + 18
// words (stage 1 also writes the cleaned U^n E that stages 2 and 3 read).
//
// Work decomposition: a block owns a TX x TY column of cells and marches
// along z over `kchunk` planes.  Planes live in LDS ring buffers:
//   raw E, B     with a 2-cell halo in x and y (the clean of the 1-cell halo
//                needs its own neighbours),
//   cleaned E    with a 1-cell halo (the stencil of rk3Push),
//   rho          on the (TX+1) x (TY+1) nodes the currents need.
// Each ring keeps only the planes its readers still need: ex, ey raw and
// cleaned 2, ez raw and cleaned 3, B 3 (bx, by of plane k+2 are stored
// only after the stencil of plane k has read plane k-1), rho 2.
//
// Per iteration (output plane k): load raw plane k+2, clean plane k+1 on
// the halo, rho on plane k+1, then rk3Push/rk3Update on plane k.  The
// per-cell arithmetic is tff_cell.h's, so results equal the unfused
// kernels bit for bit.
//
// Ghost cells: with every axis periodic (`wrap`), halo reads wrap around the
// box and no ghost cell is read or written.  Otherwise ghosts are stored:
// raw reads take the stored ghosts, and the cleaned value of a ghost cell
// comes from `ge`, which the boundary passes fill after each stage.
#pragma once
#include "tff_kernels.h"

namespace tff {

template <int TX_, int TY_>
struct Tile {
  static constexpr int TX = TX_, TY = TY_;
  static constexpr int W2 = TX + 4, H2 = TY + 4, N2 = W2 * H2;
  static constexpr int W1 = TX + 2, H1 = TY + 2, N1 = W1 * H1;
  static constexpr int WN = TX + 1, HN = TY + 1, NN = WN * HN;
  static constexpr int NT = TX * TY;
  static constexpr int L2PT = (N2 + NT - 1) / NT;
  template <class S, class T>
  static constexpr size_t lds_bytes() {
    return sizeof(S) * (size_t)(16 * N2 + 7 * N1) + sizeof(T) * (size_t)(2 * NN);
  }
};

// Field buffers are passed as one base pointer each: component c of a
// buffer starts g.cs elements after component c-1 (fewer kernel arguments
// held in scalar registers).
template <class S, class P>
struct StageArgs {
  Geom g;
  const S* ye;  // raw input E
  const S* yb;  // raw input B
  const S* ge;  // cleaned E at ghost cells (stored-ghost mode)
  const S* ue;  // U^n E for stages 2 and 3
  const S* ub;  // U^n B
  const S* t0;  // B0 curl bracket
  S* oe;        // raw output E
  S* ob;        // raw output B
  S* u0e;       // stage 1: receives the cleaned E of the input (or null)
  P cc, c1, c2, c3;
  Absorb<P> ab;
  int has_t0, stage1, clean_in, guard, wrap, kchunk;
  int nbx, nby, nblocks;  // tile grid and its size
};

__device__ __forceinline__ int wrap_index(int v, int n) {
  if (v < 0) v += n;
  while (v >= n) v -= n;
  return v;
}

__device__ __forceinline__ int s2(int p) { return p & 1; }
__device__ __forceinline__ int s3(int p) { return (p + 6) % 3; }

// One tile (tile_id in x-fastest, then y, then z-chunk order) of a stage.
template <class PR, class TL>
__device__ __forceinline__ void stage_tile(const StageArgs<typename PR::S, typename PR::P>& A, int tile_id,
                                           unsigned char* smem_raw) {
  using S = typename PR::S;
  using T = typename PR::T;
  using P = typename PR::P;
  constexpr int TX = TL::TX, TY = TL::TY, W2 = TL::W2, N2 = TL::N2, W1 = TL::W1, N1 = TL::N1;
  constexpr int WN = TL::WN, NN = TL::NN, NT = TL::NT, L2PT = TL::L2PT;
  S* rex = (S*)smem_raw;         // [2][N2]
  S* rey = rex + 2 * N2;         // [2][N2]
  S* rez = rey + 2 * N2;         // [3][N2]
  S* rbx = rez + 3 * N2;         // [3][N2]
  S* rby = rbx + 3 * N2;         // [3][N2]
  S* rbz = rby + 3 * N2;         // [3][N2]
  S* cex = rbz + 3 * N2;         // [2][N1]
  S* cey = cex + 2 * N1;         // [2][N1]
  S* cez = cey + 2 * N1;         // [3][N1]
  T* rho = (T*)(cez + 3 * N1);   // [2][NN]

  const Geom& g = A.g;
  const int64_t cs = g.cs;
  const int tid = threadIdx.x;
  const int bxi = tile_id % A.nbx;
  const int byi = (tile_id / A.nbx) % A.nby;
  const int bzi = tile_id / (A.nbx * A.nby);
  const int i0 = bxi * TX, j0 = byi * TY;
  const int k0 = bzi * A.kchunk;
  const int k1 = min(k0 + A.kchunk, g.nz);
  // tiles whose 2-cell halo stays inside the interior need no wrapping
  const bool inner_xy = i0 >= NG && i0 + TX + NG <= g.nx && j0 >= NG && j0 + TY + NG <= g.ny;

  auto addr = [&](int gi, int gj, int gk) -> int64_t {
    if (A.wrap) {
      if (!inner_xy) {
        gi = wrap_index(gi, g.nx);
        gj = wrap_index(gj, g.ny);
      }
      gk = wrap_index(gk, g.nz);
    } else {
      gi = min(gi, g.nx + NG - 1);
      gj = min(gj, g.ny + NG - 1);
    }
    return g.at(gi, gj, gk);
  };

  S reg[L2PT][6];
  auto load_plane = [&](int p) {
#pragma unroll
    for (int m = 0; m < L2PT; ++m) {
      const int c = tid + m * NT;
      if (c < N2) {
        const int lx = c % W2, ly = c / W2;
        const int64_t a = addr(i0 - 2 + lx, j0 - 2 + ly, p);
        reg[m][0] = A.ye[a];
        reg[m][1] = A.ye[a + cs];
        reg[m][2] = A.ye[a + 2 * cs];
        reg[m][3] = A.yb[a];
        reg[m][4] = A.yb[a + cs];
        reg[m][5] = A.yb[a + 2 * cs];
      }
    }
  };
  auto store_plane = [&](int p, bool with_bxy) {
#pragma unroll
    for (int m = 0; m < L2PT; ++m) {
      const int c = tid + m * NT;
      if (c < N2) {
        rex[s2(p) * N2 + c] = reg[m][0];
        rey[s2(p) * N2 + c] = reg[m][1];
        rez[s3(p) * N2 + c] = reg[m][2];
        rbz[s3(p) * N2 + c] = reg[m][5];
        if (with_bxy) {
          rbx[s3(p) * N2 + c] = reg[m][3];
          rby[s3(p) * N2 + c] = reg[m][4];
        }
      }
    }
  };
  auto store_bxy = [&](int p) {
#pragma unroll
    for (int m = 0; m < L2PT; ++m) {
      const int c = tid + m * NT;
      if (c < N2) {
        rbx[s3(p) * N2 + c] = reg[m][3];
        rby[s3(p) * N2 + c] = reg[m][4];
      }
    }
  };

  // cleanEpar of plane q on the 1-cell halo
  auto clean_plane = [&](int q) {
    for (int c = tid; c < N1; c += NT) {
      const int mx = c % W1, my = c / W1;
      const int lc = (my + 1) * W2 + (mx + 1);  // same cell in the raw plane
      const int gi = i0 - 1 + mx, gj = j0 - 1 + my;
      S ox, oy, oz;
      const bool ghost = !A.wrap && (gi < 0 || gi >= g.nx || gj < 0 || gj >= g.ny || q < 0 || q >= g.nz);
      if (!A.clean_in) {
        ox = rex[s2(q) * N2 + lc];
        oy = rey[s2(q) * N2 + lc];
        oz = rez[s3(q) * N2 + lc];
      } else if (ghost) {
        const int64_t a = g.at(min(gi, g.nx + NG - 1), min(gj, g.ny + NG - 1), q);
        ox = A.ge[a];
        oy = A.ge[a + cs];
        oz = A.ge[a + 2 * cs];
      } else {
        auto EX = [&](int a, int b, int d) { return (T)rex[s2(q + d) * N2 + lc + a + b * W2]; };
        auto EY = [&](int a, int b, int d) { return (T)rey[s2(q + d) * N2 + lc + a + b * W2]; };
        auto EZ = [&](int a, int b, int d) { return (T)rez[s3(q + d) * N2 + lc + a + b * W2]; };
        auto BXF = [&](int a, int b, int d) { return (T)rbx[s3(q + d) * N2 + lc + a + b * W2]; };
        auto BYF = [&](int a, int b, int d) { return (T)rby[s3(q + d) * N2 + lc + a + b * W2]; };
        auto BZF = [&](int a, int b, int d) { return (T)rbz[s3(q + d) * N2 + lc + a + b * W2]; };
        P px, py, pz;
        clean_cell<PR>(EX, EY, EZ, BXF, BYF, BZF, A.guard, px, py, pz);
        ox = (S)px;
        oy = (S)py;
        oz = (S)pz;
      }
      cex[s2(q) * N1 + c] = ox;
      cey[s2(q) * N1 + c] = oy;
      cez[s3(q) * N1 + c] = oz;
    }
  };

  // computeRho of plane q on the nodes of the tile and its +x, +y edge
  auto rho_plane = [&](int q) {
    for (int c = tid; c < NN; c += NT) {
      const int nx_ = c % WN, ny_ = c / WN;
      const int lc = (ny_ + 1) * W1 + (nx_ + 1);
      auto EX = [&](int a, int b, int d) { return (T)cex[s2(q + d) * N1 + lc + a + b * W1]; };
      auto EY = [&](int a, int b, int d) { return (T)cey[s2(q + d) * N1 + lc + a + b * W1]; };
      auto EZ = [&](int a, int b, int d) { return (T)cez[s3(q + d) * N1 + lc + a + b * W1]; };
      rho[s2(q) * NN + c] = rho_node<T>(EX, EY, EZ, 0, 0, 0);
    }
  };

  // prologue: planes k0-2 .. k0+1 raw, k0-1 and k0 cleaned, rho at k0
  load_plane(k0 - 2);
  store_plane(k0 - 2, true);
  load_plane(k0 - 1);
  store_plane(k0 - 1, true);
  load_plane(k0);
  store_plane(k0, true);
  __syncthreads();
  clean_plane(k0 - 1);
  __syncthreads();
  load_plane(k0 + 1);
  store_plane(k0 + 1, true);
  __syncthreads();
  clean_plane(k0);
  __syncthreads();
  rho_plane(k0);

  const int tx = tid % TX, ty = tid / TX;
  const int gi = i0 + tx, gj = j0 + ty;
  const bool active = gi < g.nx && gj < g.ny;
  const int lc1 = (ty + 1) * W1 + (tx + 1);
  const int lc2 = (ty + 2) * W2 + (tx + 2);
  const int lcn = ty * WN + tx;

  for (int k = k0; k < k1; ++k) {
    load_plane(k + 2);
    store_plane(k + 2, false);
    __syncthreads();
    clean_plane(k + 1);
    __syncthreads();
    rho_plane(k + 1);
    __syncthreads();
    if (active) {
      auto EX = [&](int a, int b, int d) { return (T)cex[s2(k + d) * N1 + lc1 + a + b * W1]; };
      auto EY = [&](int a, int b, int d) { return (T)cey[s2(k + d) * N1 + lc1 + a + b * W1]; };
      auto EZ = [&](int a, int b, int d) { return (T)cez[s3(k + d) * N1 + lc1 + a + b * W1]; };
      auto BXF = [&](int a, int b, int d) { return (T)rbx[s3(k + d) * N2 + lc2 + a + b * W2]; };
      auto BYF = [&](int a, int b, int d) { return (T)rby[s3(k + d) * N2 + lc2 + a + b * W2]; };
      auto BZF = [&](int a, int b, int d) { return (T)rbz[s3(k + d) * N2 + lc2 + a + b * W2]; };
      auto RHO = [&](int a, int b, int d) { return rho[s2(k + d) * NN + lcn + a + b * WN]; };
      const int64_t c = g.at(gi, gj, k);
      const T t0x = A.has_t0 ? (T)A.t0[c] : (T)0;
      const T t0y = A.has_t0 ? (T)A.t0[c + cs] : (T)0;
      const T t0z = A.has_t0 ? (T)A.t0[c + 2 * cs] : (T)0;
      P d[6];
      rhs_cell<PR>(EX, EY, EZ, BXF, BYF, BZF, RHO, A.cc, t0x, t0y, t0z, d);
      P v[6] = {(P)cex[s2(k) * N1 + lc1], (P)cey[s2(k) * N1 + lc1], (P)cez[s3(k) * N1 + lc1],
                (P)rbx[s3(k) * N2 + lc2], (P)rby[s3(k) * N2 + lc2], (P)rbz[s3(k) * N2 + lc2]};
      P u[6];
      if (A.stage1) {
        for (int m = 0; m < 6; ++m) u[m] = v[m];
        if (A.u0e) {
          A.u0e[c] = cex[s2(k) * N1 + lc1];
          A.u0e[c + cs] = cey[s2(k) * N1 + lc1];
          A.u0e[c + 2 * cs] = cez[s3(k) * N1 + lc1];
        }
      } else {
        u[0] = (P)A.ue[c]; u[1] = (P)A.ue[c + cs]; u[2] = (P)A.ue[c + 2 * cs];
        u[3] = (P)A.ub[c]; u[4] = (P)A.ub[c + cs]; u[5] = (P)A.ub[c + 2 * cs];
      }
      update_cell<PR>(v, u, d, A.c1, A.c2, A.c3, A.ab, gi, gj, k);
      A.oe[c] = (S)v[0];
      A.oe[c + cs] = (S)v[1];
      A.oe[c + 2 * cs] = (S)v[2];
      A.ob[c] = (S)v[3];
      A.ob[c + cs] = (S)v[4];
      A.ob[c + 2 * cs] = (S)v[5];
    }
    __syncthreads();
    store_bxy(k + 2);
  }
}

template <class PR, class TL>
__global__ void __launch_bounds__(TL::NT) k_stage(StageArgs<typename PR::S, typename PR::P> A) {
  extern __shared__ unsigned char smem_raw[];
  // Workgroups are dealt to the 8 XCDs round-robin; give each XCD one
  // contiguous run of tiles (x fastest, then y, then z-chunk) so that the
  // halos neighbouring tiles share are read through the same L2.
  const int per = (A.nblocks + 7) / 8;
  const int tile_id = (int)(blockIdx.x % 8) * per + (int)(blockIdx.x / 8);
  if (tile_id >= A.nblocks) return;
  stage_tile<PR, TL>(A, tile_id, smem_raw);
}

// ---------------------------------------------------------------------------
// Boundary passes of the stored-ghost mode, each over the ghost shell.

// The ghost cells of a new raw stage output start as the previous stage's
// cleaned fields there (rk3Update leaves ghosts alone): E from `ge`, B from
// the previous raw B.
template <class S>
__global__ void k_carry(Geom g, V3<const S> ge, V3<const S> pb, V3<S> oe, V3<S> ob, int64_t nshell) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nshell) return;
  int i, j, k;
  shell_cell(g, t, i, j, k);
  const int64_t c = g.at(i, j, k);
  oe.x[c] = ge.x[c];
  oe.y[c] = ge.y[c];
  oe.z[c] = ge.z[c];
  if (ob.x) {
    ob.x[c] = pb.x[c];
    ob.y[c] = pb.y[c];
    ob.z[c] = pb.z[c];
  }
}

// Cleaned E at every ghost cell of a raw stage output (after its hook and
// ghost fill): a ghost that exchangeFields fills holds the cleaned value of
// the wrapped interior cell, every other ghost keeps its raw value.
template <class PR>
__global__ void k_ghost_clean(Geom g, int px_, int py_, int pz_, V3<const typename PR::S> ye,
                              V3<const typename PR::S> yb, V3<typename PR::S> ge, int guard,
                              int64_t nshell) {
  using S = typename PR::S;
  using T = typename PR::T;
  using P = typename PR::P;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nshell) return;
  int i, j, k;
  shell_cell(g, t, i, j, k);
  const int64_t d = g.at(i, j, k);
  const bool ox_ = i < 0 || i >= g.nx, oy_ = j < 0 || j >= g.ny, oz_ = k < 0 || k >= g.nz;
  if ((ox_ && !px_) || (oy_ && !py_) || (oz_ && !pz_)) {
    ge.x[d] = ye.x[d];
    ge.y[d] = ye.y[d];
    ge.z[d] = ye.z[d];
    return;
  }
  const int wi = i < 0 ? i + g.nx : (i >= g.nx ? i - g.nx : i);
  const int wj = j < 0 ? j + g.ny : (j >= g.ny ? j - g.ny : j);
  const int wk = k < 0 ? k + g.nz : (k >= g.nz ? k - g.nz : k);
  const int64_t c = g.at(wi, wj, wk), sy = g.px, sz = g.plane;
  auto off = [=](int a, int b, int dd) { return c + a + b * sy + dd * sz; };
  auto EX = [&](int a, int b, int dd) { return (T)ye.x[off(a, b, dd)]; };
  auto EY = [&](int a, int b, int dd) { return (T)ye.y[off(a, b, dd)]; };
  auto EZ = [&](int a, int b, int dd) { return (T)ye.z[off(a, b, dd)]; };
  auto BXF = [&](int a, int b, int dd) { return (T)yb.x[off(a, b, dd)]; };
  auto BYF = [&](int a, int b, int dd) { return (T)yb.y[off(a, b, dd)]; };
  auto BZF = [&](int a, int b, int dd) { return (T)yb.z[off(a, b, dd)]; };
  P px, py, pz;
  clean_cell<PR>(EX, EY, EZ, BXF, BYF, BZF, guard, px, py, pz);
  ge.x[d] = (S)px;
  ge.y[d] = (S)py;
  ge.z[d] = (S)pz;
}

}  // namespace tff
