// tff_batch.h -- kernels of the batched path: many independent meshblocks
// (a parameter scan with uneven grid sizes) advanced together.
//
// Per RK stage one persistent launch of the fused stage kernel works
// through a device-side list of (instance, tile) items, so small and large
// instances share the machine without per-instance launches.
// Instances with stored ghosts then get their ghost cells in two batched
// shell passes: the new ghost values (what rk3Update leaves, the floor/top
// hook of user_loop.F90:101-164, and exchangeFields' periodic fill, applied
// in the reference's order) and the cleaned E of every ghost cell.  A
// driven loop's hook also writes its interior floor plane: one more batched
// pass, before the ghost values.
#pragma once
#include "tff_fused.h"

namespace tff {

// one instance at one stage
template <class S, class P>
struct BatchStage {
  StageArgs<S, P> a;  // fused stage kernel arguments (geometry, buffers)
  const S* ge_in;     // cleaned E at the ghosts of the stage input
  const S* prev_b;    // raw B of the stage input
  S* ge_out;          // receives the cleaned E at the ghosts of the output
};

// per-instance constants of the ghost passes
struct BatchInfo {
  int per[3];         // periodic axes
  int bc_kind;        // TFF_BC_NONE (0) or TFF_BC_LOOP (1)
  const double* bcp;  // hook parameters (device)
  int64_t nbuild;     // ghost cells of the build pass: the whole shell, or
                      // the z-ghost planes of a WRAP_MIXED instance
};

// index of the last entry of base[0..n) that is <= t
__device__ __forceinline__ int find_item(const int64_t* base, int n, int64_t t) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (base[mid] <= t) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// The work queue of k_batch_stage.  The (instance, tile) list, in instance
// order and x-fastest tile order within an instance, is cut into 8
// contiguous segments of about equal estimated cost, one per XCD.
// Workgroups are dealt to the XCDs round-robin (blockIdx % 8), so when the
// grid is a multiple of 8 the workgroups of XCD x take the items of segment
// x one at a time (neighbouring tiles, whose halo rows the XCD's L2 then
// serves once), and a workgroup whose segment is exhausted takes items from
// the next segments in turn.  Otherwise one queue holds every item.
//
// ctr[0..7] count the items taken from each segment, ctr[8] the workgroups
// that have finished; the last workgroup to finish clears all nine, so each
// launch finds them zero (the batch clears them once when it builds its
// lists) and a captured graph needs no memset node.
struct BatchQueue {
  int64_t seg[9];  // segment x is items seg[x] .. seg[x+1]-1
};
constexpr int BATCH_QWORDS = 9;

template <class PR, class TL>
__global__ void __launch_bounds__(TL::NT) __attribute__((amdgpu_waves_per_eu(sizeof(typename PR::S) == 4 ? 4 : 2)))
    k_batch_stage(const BatchStage<typename PR::S, typename PR::P>* __restrict__ st,
                  const int64_t* __restrict__ tile_base, int ninst, BatchQueue qd, unsigned* ctr) {
  extern __shared__ unsigned char smem_raw[];
  __shared__ long long s_item;
  const int nq = (gridDim.x & 7) == 0 ? 8 : 1;
  int q = nq == 8 ? (int)(blockIdx.x & 7) : 0;  // thread 0: the segment it draws from
  int left = nq;                                 // thread 0: segments not yet found empty
  for (;;) {
    __syncthreads();  // the previous item has finished with LDS and s_item
    if (threadIdx.x == 0) {
      long long item = -1;
      while (left > 0) {
        const int64_t t = qd.seg[q] + (int64_t)atomicAdd(&ctr[q], 1u);
        if (t < qd.seg[q + 1]) {
          item = t;
          break;
        }
        --left;
        q = q + 1 == nq ? 0 : q + 1;
      }
      s_item = item;
    }
    __syncthreads();
    const long long t = s_item;
    if (t < 0) break;
    const int inst = find_item(tile_base, ninst, t);
    stage_tile<PR, TL, PR::null_fp64, WRAP_RUNTIME>(st[inst].a, (int)(t - tile_base[inst]), smem_raw);
  }
  if (threadIdx.x == 0) {
    __threadfence();  // every draw of this workgroup precedes its count below
    if (atomicAdd(&ctr[8], 1u) == gridDim.x - 1) {
      for (int x = 0; x < BATCH_QWORDS; ++x) atomicExch(&ctr[x], 0u);
    }
  }
}

// The planes a FIX k_batch_stage launch left (FP32-bulk): every instance's
// tile words are one array, `mask`, in work-list order.
template <class PR, class TL>
__global__ void __launch_bounds__(TL::NT) __attribute__((amdgpu_waves_per_eu(2)))
    k_batch_stage_fix(const BatchStage<typename PR::S, typename PR::P>* __restrict__ st,
                      const int64_t* __restrict__ tile_base, int ninst, int64_t total,
                      const unsigned long long* __restrict__ mask) {
  extern __shared__ unsigned char smem_raw[];
  for (int64_t t = blockIdx.x; t < total; t += gridDim.x) {
    const unsigned long long m = mask[t];
    if (m == 0ull) continue;
    const int inst = find_item(tile_base, ninst, t);
    fix_tile<PR, TL, WRAP_RUNTIME>(st[inst].a, (int)(t - tile_base[inst]), m, smem_raw);
  }
}

// New ghost values of every stored-ghost instance's stage output.
template <class PR>
__global__ void k_batch_ghost_build(const BatchStage<typename PR::S, typename PR::P>* st,
                                    const BatchInfo* info, const int64_t* shell_base, const int* shell_inst,
                                    int nghost, int64_t total, const int* step_ptr) {
  using S = typename PR::S;
  using P = typename PR::P;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= total) return;
  // each instance's list starts on a block boundary: one lookup per block
  const int q = find_item(shell_base, nghost, (int64_t)blockIdx.x * blockDim.x);
  const int inst = shell_inst[q];
  const BatchStage<S, P>& b = st[inst];
  const BatchInfo& in = info[inst];
  const Geom& g = b.a.g;
  const int64_t cs = g.cs;
  int i, j, k;
  if (t - shell_base[q] >= in.nbuild || !shell_cell(g, t - shell_base[q], i, j, k)) return;
  const int64_t d = g.at(i, j, k);
  S* oe = b.a.oe;
  S* ob = b.a.ob;
  const bool ox = i < 0 || i >= g.nx, oy = j < 0 || j >= g.ny, oz = k < 0 || k >= g.nz;
  if (!((ox && !in.per[0]) || (oy && !in.per[1]) || (oz && !in.per[2]))) {
    // exchangeFields: the wrapped interior value (the hook never writes
    // the interior, so this is the value after the hook)
    const int wi = i < 0 ? i + g.nx : (i >= g.nx ? i - g.nx : i);
    const int wj = j < 0 ? j + g.ny : (j >= g.ny ? j - g.ny : j);
    const int wk = k < 0 ? k + g.nz : (k >= g.nz ? k - g.nz : k);
    const int64_t s = g.at(wi, wj, wk);
    for (int c = 0; c < 3; ++c) {
      oe[d + c * cs] = oe[s + c * cs];
      ob[d + c * cs] = ob[s + c * cs];
    }
    return;
  }
  // carried from the stage input
  S e3[3], b3[3];
  for (int c = 0; c < 3; ++c) {
    e3[c] = b.ge_in[d + c * cs];
    b3[c] = b.prev_b[d + c * cs];
  }
  const int mode = in.bc_kind == 1 ? (int)in.bcp[8] : -1;
  if (mode >= 0 && (oz || (mode == 2 && k == 0))) {
    // user_loop floor/top (ghost-inclusive planes): ex, ey, bz
    if (mode != 0 && k <= 0) {
      // the footpoint driver (and, in mode 2, the ghost columns of the
      // floor plane, which no periodic fill reaches)
      P ex, ey, bz;
      loop_driver<P>(in.bcp, b.a.cc, *step_ptr, i, j, k, ex, ey, bz);
      e3[0] = (S)ex;
      e3[1] = (S)ey;
      if (k < 0) b3[2] = (S)bz;
    } else if (mode == 2) {
      // conducting top
      e3[0] = (S)0;
      e3[1] = (S)0;
      b3[2] = (S)loop_top_bz<P>(in.bcp, i, j, k);
    } else {
      // copy of the k = 0 (or nz-1) value the column holds before the fill:
      // the new interior value, or for a ghost column the value carried
      // from the stage input
      const int ks = k < 0 ? 0 : g.nz - 1;
      const int64_t s = g.at(i, j, ks);
      if (!ox && !oy) {
        e3[0] = oe[s];
        e3[1] = oe[s + cs];
        b3[2] = ob[s + 2 * cs];
      } else {
        e3[0] = b.ge_in[s];
        e3[1] = b.ge_in[s + cs];
        b3[2] = b.prev_b[s + 2 * cs];
      }
    }
  }
  for (int c = 0; c < 3; ++c) {
    oe[d + c * cs] = e3[c];
    ob[d + c * cs] = b3[c];
  }
}

// The interior floor plane of every driven-loop instance (mode 2): the
// footpoint driver's ex, ey at k = 0 (bz there follows the update), written
// before the ghost pass so that periodic ghost columns copy the driven
// values.
template <class PR>
__global__ void k_batch_floor(const BatchStage<typename PR::S, typename PR::P>* st, const BatchInfo* info,
                              const int64_t* floor_base, const int* floor_inst, int nfloor, int64_t total,
                              const int* step_ptr) {
  using S = typename PR::S;
  using P = typename PR::P;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= total) return;
  const int q = find_item(floor_base, nfloor, (int64_t)blockIdx.x * blockDim.x);
  const int inst = floor_inst[q];
  const BatchStage<S, P>& b = st[inst];
  const Geom& g = b.a.g;
  const uint32_t r = (uint32_t)(t - floor_base[q]);
  if (r >= (uint32_t)g.nx * (uint32_t)g.ny) return;
  const int i = (int)(r % (uint32_t)g.nx), j = (int)(r / (uint32_t)g.nx);
  P ex, ey, bz;
  loop_driver<P>(info[inst].bcp, b.a.cc, *step_ptr, i, j, 0, ex, ey, bz);
  const int64_t c = g.at(i, j, 0);
  b.a.oe[c] = (S)ex;
  b.a.oe[c + g.cs] = (S)ey;
}

// Cleaned E at every ghost cell of the stage output (k_ghost_clean for the
// batch).
template <class PR>
__global__ void k_batch_ghost_clean(const BatchStage<typename PR::S, typename PR::P>* st,
                                    const BatchInfo* info, const int64_t* shell_base, const int* shell_inst,
                                    int nghost, int64_t total) {
  using S = typename PR::S;
  using T = typename PR::T;
  using P = typename PR::P;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= total) return;
  const int q = find_item(shell_base, nghost, (int64_t)blockIdx.x * blockDim.x);
  const int inst = shell_inst[q];
  const BatchStage<S, P>& b = st[inst];
  const BatchInfo& in = info[inst];
  const Geom& g = b.a.g;
  const int64_t cs = g.cs;
  int i, j, k;
  if (!shell_cell(g, t - shell_base[q], i, j, k)) return;
  const int64_t d = g.at(i, j, k);
  const S* ye = b.a.oe;
  const S* yb = b.a.ob;
  const bool ox_ = i < 0 || i >= g.nx, oy_ = j < 0 || j >= g.ny, oz_ = k < 0 || k >= g.nz;
  if ((ox_ && !in.per[0]) || (oy_ && !in.per[1]) || (oz_ && !in.per[2])) {
    for (int c = 0; c < 3; ++c) b.ge_out[d + c * cs] = ye[d + c * cs];
    return;
  }
  const int wi = i < 0 ? i + g.nx : (i >= g.nx ? i - g.nx : i);
  const int wj = j < 0 ? j + g.ny : (j >= g.ny ? j - g.ny : j);
  const int wk = k < 0 ? k + g.nz : (k >= g.nz ? k - g.nz : k);
  const int64_t c0 = g.at(wi, wj, wk), sy = g.px, sz = g.plane;
  auto off = [=](int a, int bb, int dd) { return c0 + a + bb * sy + dd * sz; };
  auto EX = [&](int a, int bb, int dd) { return (T)ye[off(a, bb, dd)]; };
  auto EY = [&](int a, int bb, int dd) { return (T)ye[off(a, bb, dd) + cs]; };
  auto EZ = [&](int a, int bb, int dd) { return (T)ye[off(a, bb, dd) + 2 * cs]; };
  auto BXF = [&](int a, int bb, int dd) { return (T)yb[off(a, bb, dd)]; };
  auto BYF = [&](int a, int bb, int dd) { return (T)yb[off(a, bb, dd) + cs]; };
  auto BZF = [&](int a, int bb, int dd) { return (T)yb[off(a, bb, dd) + 2 * cs]; };
  P px, py, pz;
  clean_cell<PR>(EX, EY, EZ, BXF, BYF, BZF, b.a.guard, px, py, pz);
  b.ge_out[d] = (S)px;
  b.ge_out[d + cs] = (S)py;
  b.ge_out[d + 2 * cs] = (S)pz;
}

// Cleaned E of the x/y ghost ring of planes 0 and nz-1 of every WRAP_MIXED
// instance (k_edge_clean for the batch).
template <class PR>
__global__ void k_batch_edge_clean(const BatchStage<typename PR::S, typename PR::P>* st, const int64_t* edge_base,
                                   const int* edge_inst, int nedge, int64_t total) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= total) return;
  const int q = find_item(edge_base, nedge, (int64_t)blockIdx.x * blockDim.x);
  const auto& b = st[edge_inst[q]];
  const int r = (int)(t - edge_base[q]);
  if (r >= edge_ring_cells(b.a.g)) return;
  edge_clean_at<PR>(b.a.g, b.a.oe, b.a.ob, b.ge_out, b.a.guard, r);
}

}  // namespace tff
