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
//   cleaned E    with a 1-cell halo (the stencil of rk3Push).
// Each ring keeps the planes its readers still need while the next raw
// plane is stored: ex, ey raw and cleaned 2, ez raw and cleaned 3, bz 3,
// bx and by 4 (the curl of plane k reads plane k-1).
//
// Per iteration (output plane k), two barriers (three in the fix-up kernel,
// see Ring): store raw plane k+2 (loaded into registers during the previous
// iteration) and issue the loads of plane k+3 and of U^n at plane k, so
// that their latency overlaps this iteration's work | clean plane k+1 on the
// tile and its halo | rk3Push / rk3Update on plane k, with rho evaluated at
// the four nodes each cell's currents read.  The per-cell arithmetic is
// tff_cell.h's, so results equal the unfused kernels bit for bit.
//
// FP32-bulk: an output plane whose stencil reads a near-null cell is not
// computed here but marked for k_stage_fix, which recomputes it with the
// FP64 redo; the redo's code would otherwise set this kernel's register
// budget (see stage_run).
//
// Addressing: every thread precomputes the in-plane (x-y) byte offsets of
// the cells it loads or writes once per tile; per plane only a uniform base
// pointer moves (each access is that base plus the lane's 32-bit offset).
//
// Ghost cells: with every axis periodic (`wrap`), halo reads wrap around the
// box and no ghost cell is read or written.  Otherwise ghosts are stored:
// raw reads take the stored ghosts, and the cleaned value of a ghost cell
// comes from `ge`, which the boundary passes fill after each stage.
#pragma once
#include <type_traits>

#include "tff_kernels.h"

namespace tff {

template <int TX_, int TY_>
struct Tile {
  static constexpr int TX = TX_, TY = TY_;
  static constexpr int W2 = TX + 4, H2 = TY + 4, N2 = W2 * H2;
  static constexpr int W1 = TX + 2, H1 = TY + 2, N1 = W1 * H1;
  static constexpr int NT = TX * TY;
  static constexpr int L2PT = (N2 + NT - 1) / NT;
};

// Plane rings of stage_run in LDS: how many planes of each field it keeps.
//
// The clean of plane q reads raw ex, ey at q..q+1, ez at q-1..q, bx, by at
// q-1..q and bz at q..q+1; the push of plane k reads the cleaned E at
// k-1..k+1.  TWO: the push also takes its raw-B terms (the B interpolants,
// the curl brackets and the cell's own B) from registers that the own-cell
// clean filled one plane earlier, so it reads no raw plane at all.  Raw
// plane k+3 may then be stored while other waves still push plane k, and a
// plane needs two barriers.  Otherwise (the FP32-bulk fix-up, which redoes
// near-null cells from the raw B) the push reads raw B at k-1..k+1 and a
// third barrier protects the rings.
//
// STATIC (with TWO): every ring holds 2 or 4 planes, so with the z loop
// unrolled by four (z-chunks start at multiples of 4) every ring slot is a
// compile-time constant: LDS addresses are one base register plus
// immediate offsets, and the own-cell registers alternate between two sets
// instead of being copied.  FP32 storage always takes it; FP64 takes it
// where the deeper rings leave as many workgroups per CU as rings of 3
// (64-wide and 32x16 tiles run one FP64 workgroup per CU either way), and
// otherwise keeps rings of 3 with runtime slots.
template <class PR, bool FIX, class TL>
struct Ring {
  static constexpr bool TWO = !(PR::null_fp64 && !FIX);
  // LDS bytes of TWO rings whose ez, bx, by and cleaned-ez rings hold d
  // planes (the others 2), with 8 flag words
  static constexpr size_t two_bytes(int d) {
    return sizeof(typename PR::S) * (size_t)((6 + 3 * d) * TL::N2 + (4 + d) * TL::N1) + 8 * sizeof(unsigned);
  }
  static constexpr size_t CU_BYTES = 160 * 1024;
  static constexpr bool STATIC =
      TWO && (sizeof(typename PR::S) == 4 || CU_BYTES / two_bytes(4) == CU_BYTES / two_bytes(3));
  static constexpr int EX = 2, EY = 2, EZ = STATIC ? 4 : 3;
  static constexpr int BX = STATIC ? 4 : (TWO ? 3 : 4), BY = BX, BZ = TWO ? 2 : 3;
  static constexpr int CX = 2, CY = 2, CZ = STATIC ? 4 : 3;
  static constexpr int RAW = EX + EY + EZ + BX + BY + BZ;
  static constexpr int CLEAN = CX + CY + CZ;
  // near-null flags of the planes (FP32-bulk), one word per plane slot
  static constexpr int FLAGS = TWO ? 8 : 4;
  static constexpr size_t bytes() {
    return sizeof(typename PR::S) * (size_t)(RAW * TL::N2 + CLEAN * TL::N1) + FLAGS * sizeof(unsigned);
  }
};

// LDS of a stage_run<PR, TL, FIX> workgroup
template <class PR, class TL, bool FIX>
constexpr size_t stage_lds() {
  return Ring<PR, FIX, TL>::bytes();
}
// the larger of the stage kernel's and its fix-up kernel's LDS
template <class PR, class TL>
constexpr size_t stage_lds_max() {
  return stage_lds<PR, TL, false>() > stage_lds<PR, TL, PR::null_fp64>() ? stage_lds<PR, TL, false>()
                                                                       : stage_lds<PR, TL, PR::null_fp64>();
}

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
  // near-null deferral (FP32-bulk): per tile, bit k - k0 marks output
  // plane k as left to k_stage_fix (every tile writes its word)
  unsigned long long* fixmask;
};

__device__ __forceinline__ int wrap_index(int v, int n) {
  if (v < 0) v += n;
  while (v >= n) v -= n;
  return v;
}

// element at a uniform base plus a 32-bit byte offset (the form global
// memory instructions take as scalar base + vector offset)
//
// The empty asm re-defines the offset at each access: a loop-invariant
// offset would otherwise reach the loop already widened to 64 bits, and the
// compiler then forms a 64-bit lane address per access instead of using the
// scalar-base form of the instruction.
template <class S>
__device__ __forceinline__ S ld_at(const S* base, uint32_t off) {
  asm volatile("" : "+v"(off));
  return *(const S*)((const char*)base + off);
}
// Streaming forms, for data that no other workgroup of the launch reads:
// U^n and the B0 curl (read once, at the thread's own cell) and the stage
// outputs.  The non-temporal hint keeps them from displacing the raw-plane
// lines in L2, whose halos neighbouring tiles read again a few planes later.
template <class S>
__device__ __forceinline__ S ld_once(const S* base, uint32_t off) {
  asm volatile("" : "+v"(off));
  return __builtin_nontemporal_load((const S*)((const char*)base + off));
}
template <class S>
__device__ __forceinline__ void st_at(S* base, uint32_t off, S v) {
  asm volatile("" : "+v"(off));
  __builtin_nontemporal_store(v, (S*)((char*)base + off));
}

// LDS accessor of one field around one cell: p[d+1] points at neighbour
// (-1, -1) of the cell in plane slot d = -1, 0, +1, so every stencil offset
// (a, b) in -1..1 is a small non-negative constant from one of three bases.
template <class S, int W>
struct LdsCell {
  const S* p[3];
  __device__ __forceinline__ S operator()(int a, int b, int d) const { return p[d + 1][(a + 1) + (b + 1) * W]; }
};
template <class S, int W, class SLOT>
__device__ __forceinline__ LdsCell<S, W> lds_cell(const S* ring, int plane_size, int q, int cell, SLOT slot) {
  LdsCell<S, W> r;
#pragma unroll
  for (int d = -1; d <= 1; ++d) r.p[d + 1] = ring + slot(q + d) * plane_size + cell - W - 1;
  return r;
}

// ring slot of plane p in a ring of D planes (p >= -6)
template <int D>
__device__ __forceinline__ int slot(int p) {
  if constexpr (D == 2) return p & 1;
  else if constexpr (D == 3) return (p + 6) % 3;
  else return p & 3;
}

// Ghost-cell handling of a stage launch, fixed at compile time where the
// launch knows it (A.wrap holds the same value):
//   WRAP_STORED    ghosts are stored: raw reads take them, the cleaned E of
//                  a ghost cell comes from `ge`;
//   WRAP_PERIODIC  every axis periodic: halo reads wrap around the box and
//                  no ghost is read or written;
//   WRAP_MIXED     x and y periodic, z open with a hook that reads no field:
//                  reads in the interior planes wrap in x and y (the values
//                  exchangeFields would copy there), the z-ghost planes are
//                  stored and never cleaned (their cleaned E is the raw E),
//                  and the x/y ghosts of planes 0 and nz-1, whose cleaned
//                  values depend on stored z-ghost edges, come from `ge`
//                  (k_edge_clean);
//   WRAP_RUNTIME   A.wrap decides (the batched kernel mixes instances).
constexpr int WRAP_STORED = 0, WRAP_PERIODIC = 1, WRAP_MIXED = 2, WRAP_RUNTIME = 3;

// Planes k0..k1-1 of the TX x TY column at (i0, j0) of a stage.
//
// FIX (FP32-bulk only) moves the near-null FP64 redo out of this kernel:
// the clean marks, per plane, whether any cell whose B interpolants it
// evaluates can take the redo (near_null_cell), and an output plane that
// reads such a plane (k-1, k or k+1) is skipped and marked in the tile's
// word of A.fixmask for k_stage_fix, which recomputes it with the redo.
// Skipped cells are not written, so the in-place stage 3 leaves their U^n
// for the fix-up to read.  Without the redo code the kernel fits four waves
// per SIMD.
template <class PR, class TL, bool FIX, int WM>
__device__ __forceinline__ void stage_run(const StageArgs<typename PR::S, typename PR::P>& A, int i0, int j0,
                                          int k0, int k1, int tile_id, unsigned char* smem_raw) {
  using S = typename PR::S;
  using T = typename PR::T;
  using Q = typename PR::Q;
  using P = typename PR::P;
  using RG = Ring<PR, FIX, TL>;
  constexpr int TX = TL::TX, TY = TL::TY, W2 = TL::W2, N2 = TL::N2, W1 = TL::W1, N1 = TL::N1;
  constexpr int NT = TL::NT, L2PT = TL::L2PT;
  constexpr bool TWO = RG::TWO;
  S* rex = (S*)smem_raw;
  S* rey = rex + RG::EX * N2;
  S* rez = rey + RG::EY * N2;
  S* rbx = rez + RG::EZ * N2;
  S* rby = rbx + RG::BX * N2;
  S* rbz = rby + RG::BY * N2;
  S* cex = rbz + RG::BZ * N2;
  S* cey = cex + RG::CX * N1;
  S* cez = cey + RG::CY * N1;
  unsigned* nflag = (unsigned*)(cez + RG::CZ * N1);  // plane q in slot q % FLAGS
  constexpr unsigned FM = RG::FLAGS - 1;

  const Geom& g = A.g;
  const int64_t cs = g.cs;
  const int tid = threadIdx.x;
  // Halo cell of pass m of this thread's raw-plane loads: the first pass
  // numbers the cells by thread, the later ones from the last thread down,
  // so that a partial last pass falls on the last waves and not on the
  // first ones, which also clean the halo ring.
  auto halo_cell = [&](int m) { return m == 0 ? tid : m * NT + (NT - 1 - tid); };
  const int mode = WM == WRAP_RUNTIME ? A.wrap : WM;
  const bool wrap = mode == WRAP_PERIODIC;  // every axis wraps
  const bool mixed = mode == WRAP_MIXED;    // x and y wrap, z stored
  if (FIX && tid < RG::FLAGS) nflag[tid] = 0u;

  // x-y byte offsets (from the plane origin) of this thread's halo cells
  // of the raw planes: wrapped in x and y on periodic and mixed boxes,
  // clamped to the stored ghosts otherwise; `hxs` (mixed boxes) the
  // clamped ones, for the stored z-ghost planes
  uint32_t hxy[L2PT], hxs[L2PT];
#pragma unroll
  for (int m = 0; m < L2PT; ++m) {
    const int c = min(halo_cell(m), N2 - 1);
    const int ri = i0 - 2 + c % W2, rj = j0 - 2 + c / W2;
    const int si = min(ri, g.nx + NG - 1), sj = min(rj, g.ny + NG - 1);
    hxs[m] = (uint32_t)((g.org + si + (int64_t)sj * g.px) * (int64_t)sizeof(S));
    if (wrap || mixed) {
      const int wi = wrap_index(ri, g.nx), wj = wrap_index(rj, g.ny);
      hxy[m] = (uint32_t)((g.org + wi + (int64_t)wj * g.px) * (int64_t)sizeof(S));
    } else {
      hxy[m] = hxs[m];
    }
  }
  // uniform pointer to plane p of a buffer's component 0
  auto plane_of = [&](const S* base, int p) -> const S* {
    const int gk = wrap ? wrap_index(p, g.nz) : p;
    return base + (int64_t)gk * g.plane;
  };

  // Planes are addressed twice: `q` (absolute: global rows, ghost tests,
  // flags) and `qs` (ring slots).  With STATIC rings qs is q minus a
  // multiple of 4 known at compile time, which leaves every slot the same.
  S reg[L2PT][6];
  auto load_plane = [&](int p) {
    const S* e0 = plane_of(A.ye, p);
    const S* b0 = plane_of(A.yb, p);
    const S* e1 = e0 + cs;
    const S* e2 = e0 + 2 * cs;
    const S* b1 = b0 + cs;
    const S* b2 = b0 + 2 * cs;
    const bool stored_plane = mixed && (p < 0 || p >= g.nz);
#pragma unroll
    for (int m = 0; m < L2PT; ++m) {
      if (halo_cell(m) < N2) {
        const uint32_t o = stored_plane ? hxs[m] : hxy[m];
        reg[m][0] = ld_at(e0, o);
        reg[m][1] = ld_at(e1, o);
        reg[m][2] = ld_at(e2, o);
        reg[m][3] = ld_at(b0, o);
        reg[m][4] = ld_at(b1, o);
        reg[m][5] = ld_at(b2, o);
      }
    }
  };
  auto store_plane = [&](int ps) {
#pragma unroll
    for (int m = 0; m < L2PT; ++m) {
      const int c = halo_cell(m);
      if (c < N2) {
        rex[slot<RG::EX>(ps) * N2 + c] = reg[m][0];
        rey[slot<RG::EY>(ps) * N2 + c] = reg[m][1];
        rez[slot<RG::EZ>(ps) * N2 + c] = reg[m][2];
        rbx[slot<RG::BX>(ps) * N2 + c] = reg[m][3];
        rby[slot<RG::BY>(ps) * N2 + c] = reg[m][4];
        rbz[slot<RG::BZ>(ps) * N2 + c] = reg[m][5];
      }
    }
  };

  // The B interpolants and raw |B|^2 of a thread's own cell serve both the
  // clean of that cell and, one plane later, its drift currents (they are
  // the same expressions of the same stage-input B); with TWO the own clean
  // also keeps the raw-B curl brackets and the cell's own raw B for the
  // push.  STATIC: set (q & 1) holds cleaned plane q's; otherwise the
  // clean fills set 1 and the push reads set 0.
  struct Own {
    BBar<Q> bb;
    Q bref;
    T cb[3];
    S bo[3];
  };
  Own own0{}, own1{};
  // register set i (a compile-time choice, so both stay in registers)
  auto own_set = [&](auto i) -> Own& {
    if constexpr (decltype(i)::value == 0) return own0;
    else return own1;
  };

  constexpr int RING = 2 * W1 + 2 * TY;  // halo cells of the cleaned region

  // cleanEpar of cell (mx, my) of the cleaned region of plane q; an own
  // cell (OWN) also fills `dst`
  auto clean_at = [&](int q, int qs, int mx, int my, auto own_tag, Own& dst) {
    constexpr bool OWN = decltype(own_tag)::value;
    const int c = my * W1 + mx;
    const int lc = (my + 1) * W2 + (mx + 1);  // same cell in the raw plane
    const int gi = i0 - 1 + mx, gj = j0 - 1 + my;
    const auto lex = lds_cell<S, W2>(rex, N2, qs, lc, slot<RG::EX>);
    const auto ley = lds_cell<S, W2>(rey, N2, qs, lc, slot<RG::EY>);
    const auto lez = lds_cell<S, W2>(rez, N2, qs, lc, slot<RG::EZ>);
    const auto lbx = lds_cell<S, W2>(rbx, N2, qs, lc, slot<RG::BX>);
    const auto lby = lds_cell<S, W2>(rby, N2, qs, lc, slot<RG::BY>);
    const auto lbz = lds_cell<S, W2>(rbz, N2, qs, lc, slot<RG::BZ>);
    auto EX = [&](int a, int b, int d) { return (T)lex(a, b, d); };
    auto EY = [&](int a, int b, int d) { return (T)ley(a, b, d); };
    auto EZ = [&](int a, int b, int d) { return (T)lez(a, b, d); };
    auto BXF = [&](int a, int b, int d) { return (T)lbx(a, b, d); };
    auto BYF = [&](int a, int b, int d) { return (T)lby(a, b, d); };
    auto BZF = [&](int a, int b, int d) { return (T)lbz(a, b, d); };
    S ox, oy, oz;
    const bool xyo = gi < 0 || gi >= g.nx || gj < 0 || gj >= g.ny, zo = q < 0 || q >= g.nz;
    // ghost: the cleaned value is stored in `ge`; raw: it is the raw value
    // (a mixed box's z-ghost planes, which nothing cleans)
    const bool ghost = mixed ? (!zo && xyo && (q == 0 || q == g.nz - 1)) : (!wrap && (xyo || zo));
    const bool raw = mixed && zo;
    const bool project = A.clean_in && !ghost && !raw;
    BBar<Q> bb{};
    Q bref = (Q)0;
    if (project || OWN) {
      bb = bbar<T, Q, PR::null_fp64>(BXF, BYF, BZF);
      if (PR::null_fp64) bref = cell_b2<Q>(BXF, BYF, BZF);
      if (FIX && near_null_cell<Q>(bb, bref)) nflag[(unsigned)q & FM] = 1u;
    }
    if (!A.clean_in || raw) {
      ox = lex(0, 0, 0);
      oy = ley(0, 0, 0);
      oz = lez(0, 0, 0);
    } else if (ghost) {
      const int64_t a = g.at(min(gi, g.nx + NG - 1), min(gj, g.ny + NG - 1), q);
      ox = A.ge[a];
      oy = A.ge[a + cs];
      oz = A.ge[a + 2 * cs];
    } else {
      P px, py, pz;
      clean_cell_bb<PR, PR::null_fp64 && !FIX>(EX, EY, EZ, BXF, BYF, BZF, bb, bref, A.guard, px, py, pz);
      ox = (S)px;
      oy = (S)py;
      oz = (S)pz;
    }
    cex[slot<RG::CX>(qs) * N1 + c] = ox;
    cey[slot<RG::CY>(qs) * N1 + c] = oy;
    cez[slot<RG::CZ>(qs) * N1 + c] = oz;
    if constexpr (OWN) {
      dst.bb = bb;
      dst.bref = bref;
      if constexpr (TWO) {
        curl_b<T>(BXF, BYF, BZF, dst.cb);
        dst.bo[0] = lbx(0, 0, 0);
        dst.bo[1] = lby(0, 0, 0);
        dst.bo[2] = lbz(0, 0, 0);
      }
    }
  };
  // One component PT of the cleaned E at cell (mx, my) of the halo ring:
  // clean_at's value of that component, with only E point PT's B
  // interpolants evaluated (policies without the near-null redo, whose
  // test needs all three points)
  auto clean_one = [&](int q, int qs, int mx, int my, auto ptag) {
    constexpr int PT = decltype(ptag)::value;
    const int c = my * W1 + mx;
    const int lc = (my + 1) * W2 + (mx + 1);
    const int gi = i0 - 1 + mx, gj = j0 - 1 + my;
    const bool xyo = gi < 0 || gi >= g.nx || gj < 0 || gj >= g.ny, zo = q < 0 || q >= g.nz;
    const bool ghost = mixed ? (!zo && xyo && (q == 0 || q == g.nz - 1)) : (!wrap && (xyo || zo));
    const bool raw = mixed && zo;
    const auto lex = lds_cell<S, W2>(rex, N2, qs, lc, slot<RG::EX>);
    const auto ley = lds_cell<S, W2>(rey, N2, qs, lc, slot<RG::EY>);
    const auto lez = lds_cell<S, W2>(rez, N2, qs, lc, slot<RG::EZ>);
    S o;
    if (!A.clean_in || raw) {
      o = PT == 0 ? lex(0, 0, 0) : (PT == 1 ? ley(0, 0, 0) : lez(0, 0, 0));
    } else if (ghost) {
      const int64_t a = g.at(min(gi, g.nx + NG - 1), min(gj, g.ny + NG - 1), q);
      o = A.ge[a + PT * cs];
    } else {
      const auto lbx = lds_cell<S, W2>(rbx, N2, qs, lc, slot<RG::BX>);
      const auto lby = lds_cell<S, W2>(rby, N2, qs, lc, slot<RG::BY>);
      const auto lbz = lds_cell<S, W2>(rbz, N2, qs, lc, slot<RG::BZ>);
      auto EX = [&](int a, int b, int d) { return (T)lex(a, b, d); };
      auto EY = [&](int a, int b, int d) { return (T)ley(a, b, d); };
      auto EZ = [&](int a, int b, int d) { return (T)lez(a, b, d); };
      auto BXF = [&](int a, int b, int d) { return (T)lbx(a, b, d); };
      auto BYF = [&](int a, int b, int d) { return (T)lby(a, b, d); };
      auto BZF = [&](int a, int b, int d) { return (T)lbz(a, b, d); };
      T bx, by, bz;
      bbar_point<PT, T>(BXF, BYF, BZF, bx, by, bz);
      const Q x = bx, y = by, z = bz;
      o = (S)(P)clean_point<PT, T, Q, false>(EX, EY, EZ, x, y, z, sq3<false>(x, y, z), A.guard);
    }
    S* ce = PT == 0 ? cex : (PT == 1 ? cey : cez);
    constexpr int D = PT == 0 ? RG::CX : (PT == 1 ? RG::CY : RG::CZ);
    ce[slot<D>(qs) * N1 + c] = o;
  };
  // The halo-ring values the push reads (of plane k at k-1, k and k+1), by
  // component, in ring coordinates (mx, my) = (i + 1, j + 1):
  //   Ex  x = -1 for j = 0..TY, x = TX for j < TY, y = TY for i < TX;
  //   Ey  y = -1 for i = 0..TX, x = TX for j < TY, y = TY for i < TX;
  //   Ez  x = TX for j < TY, y = TY for i < TX.
  // SPLIT (no near-null redo) cleans just these, one component per item and
  // one component per wave: 290 items of a third of a clean each on a 64x8
  // tile, in 7 waves, against 148 whole cells in 3.
  constexpr bool SPLIT = !PR::null_fp64;
  constexpr int NRX = 2 * TY + 1 + TX, NRY = 2 * TX + TY + 1, NRZ = TX + TY;
  constexpr int NCX = (NRX + 63) / 64, NCY = (NRY + 63) / 64, NCZ = (NRZ + 63) / 64;
  constexpr int NCH = NCX + NCY + NCZ, NW = NT / 64;
  using I0 = std::integral_constant<int, 0>;
  using I1 = std::integral_constant<int, 1>;
  using I2 = std::integral_constant<int, 2>;
  auto clean_ring_split = [&](int q, int qs) {
    const int lane = tid & 63;
    for (int ch = tid >> 6; ch < NCH; ch += NW) {  // uniform per wave
      if (ch < NCX) {
        const int e = ch * 64 + lane;
        if (e < NRX) {
          const int mx = e <= TY ? 0 : (e < 2 * TY + 1 ? TX + 1 : e - 2 * TY);
          const int my = e <= TY ? e + 1 : (e < 2 * TY + 1 ? e - TY : TY + 1);
          clean_one(q, qs, mx, my, I0{});
        }
      } else if (ch < NCX + NCY) {
        const int e = (ch - NCX) * 64 + lane;
        if (e < NRY) {
          const int mx = e <= TX ? e + 1 : (e < TX + 1 + TY ? TX + 1 : e - TX - TY);
          const int my = e <= TX ? 0 : (e < TX + 1 + TY ? e - TX : TY + 1);
          clean_one(q, qs, mx, my, I1{});
        }
      } else {
        const int e = (ch - NCX - NCY) * 64 + lane;
        if (e < NRZ) {
          const int mx = e < TY ? TX + 1 : e - TY + 1;
          const int my = e < TY ? e + 1 : TY + 1;
          clean_one(q, qs, mx, my, I2{});
        }
      }
    }
  };
  // cleanEpar of plane q on the tile and its 1-cell halo: each thread its
  // own cell, then the halo ring: whole cells by the first RING threads, or
  // with SPLIT the components the push reads
  // (own_tag false: the own cell's values are not needed)
  auto clean_plane = [&](int q, int qs, auto own_tag, Own& dst) {
    // the waves that also clean the halo ring reach the next barrier last:
    // they issue first while they clean
    const bool ring_wave = SPLIT ? (tid >> 6) < NCH : (tid & ~63) < RING;
    if (ring_wave) __builtin_amdgcn_s_setprio(2);
    clean_at(q, qs, tid % TX + 1, tid / TX + 1, own_tag, dst);
    if constexpr (SPLIT) {
      clean_ring_split(q, qs);
    } else {
      for (int r = tid; r < RING; r += NT) {
        int mx, my;
        if (r < W1) {
          mx = r;
          my = 0;
        } else if (r < 2 * W1) {
          mx = r - W1;
          my = TY + 1;
        } else if (r < 2 * W1 + TY) {
          mx = 0;
          my = r - 2 * W1 + 1;
        } else {
          mx = W1 - 1;
          my = r - 2 * W1 - TY + 1;
        }
        clean_at(q, qs, mx, my, std::false_type{}, dst);
      }
    }
    if (ring_wave) __builtin_amdgcn_s_setprio(0);
  };

  // prologue: planes k0-2 .. k0+1 raw, k0-1 and k0 cleaned (k0's own-cell
  // values where the first push reads them)
  load_plane(k0 - 2);
  store_plane(k0 - 2);
  load_plane(k0 - 1);
  store_plane(k0 - 1);
  load_plane(k0);
  store_plane(k0);
  __syncthreads();
  clean_plane(k0 - 1, k0 - 1, std::false_type{}, own0);
  __syncthreads();
  load_plane(k0 + 1);
  store_plane(k0 + 1);
  __syncthreads();
  clean_plane(k0, k0, std::true_type{}, own_set(std::integral_constant<int, RG::STATIC ? 0 : 1>{}));
  __syncthreads();
  if constexpr (!RG::STATIC) own0 = own1;
  if (k0 < k1) load_plane(k0 + 2);

  const int tx = tid % TX, ty = tid / TX;
  const int gi = i0 + tx, gj = j0 + ty;
  const bool active = gi < g.nx && gj < g.ny;
  const int lc1 = (ty + 1) * W1 + (tx + 1);
  const int lc2 = (ty + 2) * W2 + (tx + 2);
  // x-y byte offset of the thread's own cell (clamped for inactive threads)
  const uint32_t oxy =
      (uint32_t)((g.org + min(gi, g.nx - 1) + (int64_t)min(gj, g.ny - 1) * g.px) * (int64_t)sizeof(S));

  unsigned long long skipped = 0ull;  // FIX: bit k - k0 for a skipped plane k
  // Output plane k; PH is k mod 4 when known at compile time (STATIC),
  // else -1.
  auto iteration = [&](int k, auto ph) {
    constexpr int PH = decltype(ph)::value;
    const int ks = PH >= 0 ? PH : k;
    // raw plane k+2 arrived during the previous iteration (with TWO, other
    // waves may still push plane k-1: it reads no raw slot)
    store_plane(ks + 2);
    // the flag slot of plane k+2 (its previous plane was last read before
    // this iteration)
    if (FIX && tid == 0) nflag[(unsigned)(k + 2) & FM] = 0u;
    const int64_t pk = (int64_t)k * g.plane;
    // U^n and the B0 curl of plane k, then raw plane k+3, in flight while
    // this iteration cleans and computes (FP64 loads U^n where it is used:
    // two more FP64 sextets in flight would cost a wave per SIMD)
    constexpr bool EARLY = sizeof(S) == 4;
    S un[6];
    T tt[3] = {(T)0, (T)0, (T)0};
    auto load_un = [&]() {
      if (!A.stage1) {
        const S* ue = A.ue + pk;
        const S* ub = A.ub + pk;
        un[0] = ld_once(ue, oxy);
        un[1] = ld_once(ue + cs, oxy);
        un[2] = ld_once(ue + 2 * cs, oxy);
        un[3] = ld_once(ub, oxy);
        un[4] = ld_once(ub + cs, oxy);
        un[5] = ld_once(ub + 2 * cs, oxy);
      }
      if (A.has_t0) {
        const S* t0 = A.t0 + pk;
        tt[0] = (T)ld_once(t0, oxy);
        tt[1] = (T)ld_once(t0 + cs, oxy);
        tt[2] = (T)ld_once(t0 + 2 * cs, oxy);
      }
    };
    if (EARLY && active) load_un();
    if (k + 1 < k1) load_plane(k + 3);
    __syncthreads();
    clean_plane(k + 1, ks + 1, std::true_type{}, own_set(std::integral_constant<int, PH >= 0 ? (PH + 1) & 1 : 1>{}));
    __syncthreads();
    const Own& cur = own_set(std::integral_constant<int, PH >= 0 ? PH & 1 : 0>{});
    bool skip = false;
    if constexpr (FIX) {
      const unsigned f = nflag[(unsigned)(k - 1) & FM] | nflag[(unsigned)k & FM] | nflag[(unsigned)(k + 1) & FM];
      skip = __builtin_amdgcn_readfirstlane(f) != 0u;
      if (skip) skipped |= 1ull << (k - k0);
    }
    if (active && !skip) {
      if (!EARLY) load_un();
      const auto lex = lds_cell<S, W1>(cex, N1, ks, lc1, slot<RG::CX>);
      const auto ley = lds_cell<S, W1>(cey, N1, ks, lc1, slot<RG::CY>);
      const auto lez = lds_cell<S, W1>(cez, N1, ks, lc1, slot<RG::CZ>);
      auto EX = [&](int a, int b, int d) { return (T)lex(a, b, d); };
      auto EY = [&](int a, int b, int d) { return (T)ley(a, b, d); };
      auto EZ = [&](int a, int b, int d) { return (T)lez(a, b, d); };
      // computeRho at the four nodes the currents read, from the cleaned
      // rings (cheaper than a separate rho pass and its barrier)
      auto RHO = [&](int a, int b, int d) { return rho_node<T>(EX, EY, EZ, a, b, d); };
      P d[6];
      P v[6];
      v[0] = (P)lex(0, 0, 0);
      v[1] = (P)ley(0, 0, 0);
      v[2] = (P)lez(0, 0, 0);
      if constexpr (TWO) {
        rhs_cell_pre<PR>(EX, EY, EZ, RHO, cur.bb, cur.bref, A.cc, tt[0], tt[1], tt[2], cur.cb, d);
        v[3] = (P)cur.bo[0];
        v[4] = (P)cur.bo[1];
        v[5] = (P)cur.bo[2];
      } else {
        const auto lbx = lds_cell<S, W2>(rbx, N2, ks, lc2, slot<RG::BX>);
        const auto lby = lds_cell<S, W2>(rby, N2, ks, lc2, slot<RG::BY>);
        const auto lbz = lds_cell<S, W2>(rbz, N2, ks, lc2, slot<RG::BZ>);
        auto BXF = [&](int a, int b, int d) { return (T)lbx(a, b, d); };
        auto BYF = [&](int a, int b, int d) { return (T)lby(a, b, d); };
        auto BZF = [&](int a, int b, int d) { return (T)lbz(a, b, d); };
        rhs_cell_bb<PR, PR::null_fp64 && !FIX>(EX, EY, EZ, BXF, BYF, BZF, RHO, cur.bb, cur.bref, A.cc, tt[0],
                                               tt[1], tt[2], d);
        v[3] = (P)lbx(0, 0, 0);
        v[4] = (P)lby(0, 0, 0);
        v[5] = (P)lbz(0, 0, 0);
      }
      P u[6];
      if (A.stage1) {
        for (int m = 0; m < 6; ++m) u[m] = v[m];
        if (A.u0e) {
          S* u0 = A.u0e + pk;
          st_at(u0, oxy, lex(0, 0, 0));
          st_at(u0 + cs, oxy, ley(0, 0, 0));
          st_at(u0 + 2 * cs, oxy, lez(0, 0, 0));
        }
      } else {
        for (int m = 0; m < 6; ++m) u[m] = (P)un[m];
      }
      update_cell<PR>(v, u, d, A.c1, A.c2, A.c3, A.ab, gi, gj, k);
      S* oe = A.oe + pk;
      S* ob = A.ob + pk;
      st_at(oe, oxy, (S)v[0]);
      st_at(oe + cs, oxy, (S)v[1]);
      st_at(oe + 2 * cs, oxy, (S)v[2]);
      st_at(ob, oxy, (S)v[3]);
      st_at(ob + cs, oxy, (S)v[4]);
      st_at(ob + 2 * cs, oxy, (S)v[5]);
    }
    if constexpr (PH < 0) own0 = own1;
    if constexpr (!TWO) __syncthreads();  // plane k's rings are free for plane k+3 / k+2
  };
  if constexpr (RG::STATIC) {
    // k0 is a multiple of 4 (z-chunks are), so plane kk + j has phase j
    for (int kk = k0; kk < k1; kk += 4) {
      iteration(kk, std::integral_constant<int, 0>{});
      if (kk + 1 < k1) iteration(kk + 1, std::integral_constant<int, 1>{});
      if (kk + 2 < k1) iteration(kk + 2, std::integral_constant<int, 2>{});
      if (kk + 3 < k1) iteration(kk + 3, std::integral_constant<int, 3>{});
    }
  } else {
    for (int k = k0; k < k1; ++k) iteration(k, std::integral_constant<int, -1>{});
  }
  if (FIX && tid == 0) A.fixmask[tile_id] = skipped;
}

// One tile (tile_id in x-fastest, then y, then z-chunk order) of a stage.
template <class PR, class TL, bool FIX, int WM>
__device__ __forceinline__ void stage_tile(const StageArgs<typename PR::S, typename PR::P>& A, int tile_id,
                                           unsigned char* smem_raw) {
  const int bxi = tile_id % A.nbx;
  const int byi = (tile_id / A.nbx) % A.nby;
  const int bzi = tile_id / (A.nbx * A.nby);
  const int k0 = bzi * A.kchunk;
  stage_run<PR, TL, FIX, WM>(A, bxi * TL::TX, byi * TL::TY, k0, min(k0 + A.kchunk, A.g.nz), tile_id, smem_raw);
}

template <class PR, class TL, int WM>
__global__ void __launch_bounds__(TL::NT) __attribute__((amdgpu_waves_per_eu(sizeof(typename PR::S) == 4 ? 4 : 2)))
    k_stage(StageArgs<typename PR::S, typename PR::P> A) {
  extern __shared__ unsigned char smem_raw[];
  // Workgroups are dealt to the 8 XCDs round-robin; give each XCD one
  // contiguous run of tiles (x fastest, then y, then z-chunk) so that the
  // halos neighbouring tiles share are read through the same L2.
  const int per = (A.nblocks + 7) / 8;
  const int tile_id = (int)(blockIdx.x % 8) * per + (int)(blockIdx.x / 8);
  if (tile_id >= A.nblocks) return;
  stage_tile<PR, TL, PR::null_fp64, WM>(A, tile_id, smem_raw);
}

// The planes of tile t that a FIX stage launch marked in `m`, with the
// near-null FP64 redo: the same per-cell code as the stage kernel of the
// unfused path, so the results are those of a kernel that redoes inline.
// Each run of consecutive marked planes is one stage_run.
template <class PR, class TL, int WM>
__device__ __forceinline__ void fix_tile(const StageArgs<typename PR::S, typename PR::P>& A, int t,
                                         unsigned long long m, unsigned char* smem_raw) {
  const int i0 = (t % A.nbx) * TL::TX;
  const int j0 = ((t / A.nbx) % A.nby) * TL::TY;
  const int k0 = (t / (A.nbx * A.nby)) * A.kchunk;
  while (m != 0ull) {
    const int a = __builtin_ctzll(m);  // first marked plane
    const unsigned long long rest = ~(m >> a);
    const int len = rest == 0ull ? 64 - a : __builtin_ctzll(rest);  // marked planes from a on
    stage_run<PR, TL, false, WM>(A, i0, j0, k0 + a, k0 + a + len, t, smem_raw);
    m = (a + len >= 64) ? 0ull : (m >> (a + len)) << (a + len);
  }
}

// After a FIX k_stage launch: block b takes tiles b, b + gridDim.x, ...
template <class PR, class TL, int WM>
__global__ void __launch_bounds__(TL::NT) __attribute__((amdgpu_waves_per_eu(2)))
    k_stage_fix(StageArgs<typename PR::S, typename PR::P> A) {
  extern __shared__ unsigned char smem_raw[];
  for (int t = (int)blockIdx.x; t < A.nblocks; t += (int)gridDim.x) {
    const unsigned long long m = A.fixmask[t];
    if (m != 0ull) fix_tile<PR, TL, WM>(A, t, m, smem_raw);
  }
}

// ---------------------------------------------------------------------------
// Boundary passes of the stored-ghost mode, each over the ghost shell.

// The ghost cells of a new raw stage output start as the previous stage's
// cleaned fields there (rk3Update leaves ghosts alone): E from `ge`, B from
// the previous raw B.  With `per` (the periodic axes) set, a ghost cell that
// the periodic fill will overwrite before anything reads it is skipped.
template <class S>
__global__ void k_carry(Geom g, V3<const S> ge, V3<const S> pb, V3<S> oe, V3<S> ob, int64_t nshell, int3 per) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nshell) return;
  int i, j, k;
  shell_cell(g, t, i, j, k);
  const bool ox = i < 0 || i >= g.nx, oy = j < 0 || j >= g.ny, oz = k < 0 || k >= g.nz;
  if (!((ox && !per.x) || (oy && !per.y) || (oz && !per.z))) return;
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

// Cleaned E of the x/y ghost cells of planes 0 and nz-1 of a WRAP_MIXED
// box (the 1-cell ring around each plane), into `ge`.  exchangeFields
// copies there the cleaned value of the wrapped interior cell, whose
// stencil reaches into the z-ghost planes: its neighbours in the interior
// planes are read wrapped in x and y (the copies exchangeFields makes),
// those in a z-ghost plane at their stored place (exchangeFields leaves the
// edges of an open axis alone).  Buffers are component-0 bases, components
// g.cs apart; t numbers the ring cells of plane 0, then of plane nz-1.
__device__ __forceinline__ int edge_ring_cells(const Geom& g) {
  return (2 * (g.nx + 2) + 2 * g.ny) * (g.nz > 1 ? 2 : 1);
}
template <class PR>
__device__ __forceinline__ void edge_clean_at(const Geom& g, const typename PR::S* ye, const typename PR::S* yb,
                                              typename PR::S* ge, int guard, int t) {
  using S = typename PR::S;
  using T = typename PR::T;
  using P = typename PR::P;
  const int ring = 2 * (g.nx + 2) + 2 * g.ny;
  const int q = t < ring ? 0 : g.nz - 1;
  const int r = t < ring ? t : t - ring;
  int i, j;
  if (r < g.nx + 2) {
    i = r - 1;
    j = -1;
  } else if (r < 2 * (g.nx + 2)) {
    i = r - (g.nx + 2) - 1;
    j = g.ny;
  } else if (r < 2 * (g.nx + 2) + g.ny) {
    i = -1;
    j = r - 2 * (g.nx + 2);
  } else {
    i = g.nx;
    j = r - 2 * (g.nx + 2) - g.ny;
  }
  const int wi = wrap_index(i, g.nx), wj = wrap_index(j, g.ny);
  const int64_t cs = g.cs;
  auto at = [&](int a, int b, int d) -> int64_t {
    const int k = q + d;
    if (k >= 0 && k < g.nz) return g.at(wrap_index(wi + a, g.nx), wrap_index(wj + b, g.ny), k);
    return g.at(wi + a, wj + b, k);
  };
  auto EX = [&](int a, int b, int d) { return (T)ye[at(a, b, d)]; };
  auto EY = [&](int a, int b, int d) { return (T)ye[at(a, b, d) + cs]; };
  auto EZ = [&](int a, int b, int d) { return (T)ye[at(a, b, d) + 2 * cs]; };
  auto BXF = [&](int a, int b, int d) { return (T)yb[at(a, b, d)]; };
  auto BYF = [&](int a, int b, int d) { return (T)yb[at(a, b, d) + cs]; };
  auto BZF = [&](int a, int b, int d) { return (T)yb[at(a, b, d) + 2 * cs]; };
  P px, py, pz;
  clean_cell<PR>(EX, EY, EZ, BXF, BYF, BZF, guard, px, py, pz);
  const int64_t d = g.at(i, j, q);
  ge[d] = (S)px;
  ge[d + cs] = (S)py;
  ge[d + 2 * cs] = (S)pz;
}

template <class PR>
__global__ void k_edge_clean(Geom g, const typename PR::S* ye, const typename PR::S* yb, typename PR::S* ge,
                             int guard) {
  const int t = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (t >= edge_ring_cells(g)) return;
  edge_clean_at<PR>(g, ye, yb, ge, guard, t);
}

}  // namespace tff
