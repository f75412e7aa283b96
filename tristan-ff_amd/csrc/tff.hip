// tff.hip -- host side of libtff.so: the C ABI declared in include/tff.h.
//
// A handle holds one meshblock in HBM as structure-of-arrays buffers (one
// per field component, layout in tff_common.h) and one HIP stream.  The
// reference's module arrays ex..bz are the handle's state buffers; each
// kernel writes a fresh buffer and the host swaps pointers, so no kernel
// ever reads what another thread of the same launch writes.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>
#include <string>
#include <type_traits>
#include <vector>

#include "tff.h"
#include "tff_batch.h"
#include "tff_output.h"
#include "tff_init.h"

using namespace tff;

namespace {

struct Vec {
  void* p[3] = {nullptr, nullptr, nullptr};
};

template <class S>
V3<S> mv3(const Vec& v) {
  return V3<S>{(S*)v.p[0], (S*)v.p[1], (S*)v.p[2]};
}
template <class S>
V3<const S> cv3(const Vec& v) {
  return V3<const S>{(const S*)v.p[0], (const S*)v.p[1], (const S*)v.p[2]};
}

}  // namespace

struct tff_handle {
  tff_config cfg{};
  size_t esz = 8;  // storage element size
  Geom g{};
  int xo = 0;           // elements before the first interior cell of a row
  int64_t nalloc = 0;   // elements per component buffer
  hipStream_t stream = nullptr;
  std::string err;

  // state (the reference's ex..bz), stage work buffers and the clean target
  Vec es, bs, w1e, w1b, w2e, w2b, ce;
  Vec en, bn;  // initRKstep copies (allocated on first use)
  Vec b0, t0;  // background field and its curl bracket
  bool t0_dirty = true;
  int has_t0 = 0;
  int* d_flag = nullptr;
  // FP32-bulk fused path: one word per stage-kernel tile marking the planes
  // it left to k_stage_fix (room for fix_cap tiles)
  unsigned long long* d_fixmask = nullptr;
  int64_t fix_cap = 0;

  int bc_kind = TFF_BC_NONE;
  double bc_host[16] = {0};
  double* d_bc = nullptr;
  void* psr_stage = nullptr;
  int3 psr_lo{0, 0, 0}, psr_n{0, 0, 0};

  int* d_step = nullptr;
  double* d_part = nullptr;
  double* d_out = nullptr;
  int red_blocks = 0;

  hipEvent_t ev[16] = {};
  // phase timers (tff_set_phase_timing): event pairs per phase, recorded
  // around the solver kernels, the user hook and the ghost exchange
  bool phases = false;
  std::vector<hipEvent_t> ppool;
  std::vector<int> pphase;  // phase of each recorded pair
  size_t pused = 0;
  double psum_ms[TFF_PHASES] = {0.0};
  bool timing = false;
  std::vector<hipEvent_t> tpool;
  size_t tused = 0;
  double tsum_ms = 0.0;
  long tcount = 0;

  // fused path: cleaned E at ghost cells, whether the state buffers hold the
  // cleaned (reference) state or the raw output of the last stage, and the
  // tile shape and z-chunk picked by the autotuner
  Vec ge;
  bool state_clean = true;
  // every buffer's ghost cells hold the same values, which no later stage
  // changes (see ghosts_static): the fused step then runs only the hook
  bool ghosts_settled = false;
  int fused = 1;
  // TFF_FP32_COMP: every field buffer carries lo words after its hi words
  // (alloc_vec doubles it); the unfused path only
  bool comp = false;
  int tile = -1;
  int kchunk = 0;
  int max_lds = 65536;

  // keyed by the state and clean E buffers and the number of steps captured
  std::map<std::tuple<void*, void*, int>, hipGraphExec_t> graphs;
  std::vector<hipGraphExec_t> retired;  // invalidated, not yet destroyed
  std::vector<void*> allocs;
  // staging of tff_output_fields (grown on demand, kept until destroy)
  void* out_buf = nullptr;
  size_t out_cap = 0;
};

namespace {

// Events around one phase of a step while phase timing is on (nothing
// otherwise); the pairs are read and summed by tff_phase_times.
struct PhaseScope {
  tff_handle* h;
  hipEvent_t e1 = nullptr;
  PhaseScope(tff_handle* hh, int phase) : h(hh) {
    if (!h->phases) return;
    while (h->ppool.size() < h->pused + 2) {
      hipEvent_t e;
      if (hipEventCreate(&e) != hipSuccess) {
        h->phases = false;
        return;
      }
      h->ppool.push_back(e);
    }
    hipEventRecord(h->ppool[h->pused], h->stream);
    e1 = h->ppool[h->pused + 1];
    h->pphase.resize(h->pused / 2 + 1);
    h->pphase[h->pused / 2] = phase;
    h->pused += 2;
  }
  ~PhaseScope() {
    if (e1) hipEventRecord(e1, h->stream);
  }
};

int fail(tff_handle* h, int code, const std::string& msg) {
  if (h) h->err = msg;
  return code;
}

#define HIPCHK(h, call)                                                                   \
  do {                                                                                    \
    hipError_t e_ = (call);                                                               \
    if (e_ != hipSuccess)                                                                 \
      return fail((h), TFF_EHIP, std::string(#call) + ": " + hipGetErrorString(e_));      \
  } while (0)

// One allocation per field buffer: the x, y and z components follow each
// other nalloc elements apart (Geom::cs).  Every transfer and fill goes on
// the handle's stream: it is a non-blocking stream, so work on the null
// stream (hipMemset, hipMemcpy) would not be ordered before its kernels.
int alloc_vec(tff_handle* h, Vec& v) {
  void* p = nullptr;
  const size_t nb = (h->comp ? 6 : 3) * (size_t)h->nalloc * h->esz;
  if (hipMalloc(&p, nb) != hipSuccess) return fail(h, TFF_ENOMEM, "hipMalloc of a field buffer failed");
  if (hipMemsetAsync(p, 0, nb, h->stream) != hipSuccess) {
    hipFree(p);
    return fail(h, TFF_EHIP, "hipMemsetAsync failed");
  }
  h->allocs.push_back(p);
  for (int c = 0; c < 3; ++c) v.p[c] = (char*)p + (size_t)c * h->nalloc * h->esz;
  return TFF_OK;
}

template <class S>
const S* base_c(const Vec& v) {
  return (const S*)v.p[0];
}
template <class S>
S* base_m(const Vec& v) {
  return (S*)v.p[0];
}

// Stop replaying the captured graphs (buffer roles or hook parameters
// changed).  They are destroyed later, where the library captures anew or
// tears down, so that entry points that only clean the state never free.
void invalidate_graphs(tff_handle* h) {
  for (auto& kv : h->graphs) h->retired.push_back(kv.second);
  h->graphs.clear();
  h->ghosts_settled = false;
}

// destroy the graphs invalidate_graphs set aside (the stream is idle or
// ordered after their last launch)
void drop_retired(tff_handle* h) {
  for (auto e : h->retired) hipGraphExecDestroy(e);
  h->retired.clear();
}

int last_launch(tff_handle* h, const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(h, TFF_EHIP, std::string(what) + ": " + hipGetErrorString(e));
  return TFF_OK;
}

dim3 box_grid(const tff_handle* h) {
  const int sx = h->g.nx + 2 * NG, sy = h->g.ny + 2 * NG, sz = h->g.nz + 2 * NG;
  return dim3((sx + BX - 1) / BX, (sy + BY - 1) / BY, sz);
}

// ---- precision-dispatched launches ---------------------------------------
template <class PR>
int launch_rk3(tff_handle* h, const Vec& xe, const Vec& xb, const Vec& ue, const Vec& ub,
               const Vec& ye, const Vec& yb, double c1, double c2, double c3) {
  using S = typename PR::S;
  PhaseScope phase(h, TFF_PHASE_SOLVER);
  using P = typename PR::P;
  Absorb<P> ab{};
  ab.on = h->cfg.boundary[0] == TFF_ABSORB;
  ab.fixed = h->cfg.absorb_fixed;
  // field_solver.F90:200-204, evaluated in the working precision
  ab.xc = (P)0.5 * (P)h->g.nx;
  ab.yc = (P)0.5 * (P)h->g.ny;
  ab.zc = (P)0.5 * (P)h->g.nz;
  P rmax = std::min(std::min(ab.xc, ab.yc), ab.zc);
  ab.rabsorb = rmax - (P)h->cfg.absorb_buff;
  ab.buff = (P)h->cfg.absorb_buff;
  // FP64 and FP32 take CC and the stage coefficients in their own precision
  // (the reference's default-REAL literals); FP32-bulk accumulates in FP64
  // with the exact FP64 values.
  const bool exact = std::is_same<S, P>::value;
  const P cc = exact ? (P)(S)h->cfg.cc : (P)h->cfg.cc;
  const P k1 = exact ? (P)(S)c1 : (P)c1, k2 = exact ? (P)(S)c2 : (P)c2, k3 = exact ? (P)(S)c3 : (P)c3;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (h->timing) {
    while (h->tpool.size() < h->tused + 2) {
      hipEvent_t e;
      // timing only: no system-scope fence (its cache writeback and
      // invalidation would slow the kernel it brackets)
      HIPCHK(h, hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
      h->tpool.push_back(e);
    }
    e0 = h->tpool[h->tused++];
    e1 = h->tpool[h->tused++];
    HIPCHK(h, hipEventRecord(e0, h->stream));
  }
  if constexpr (std::is_same<PR, PrecF32Bulk>::value && sizeof(P) == 4) {
    if (h->comp) {
      Absorb<float> af{ab.on, ab.fixed, (float)ab.xc, (float)ab.yc, (float)ab.zc, (float)ab.rabsorb,
                       (float)ab.buff};
      hipLaunchKernelGGL(k_rk3c<PR>, box_grid(h), dim3(BX, BY), 0, h->stream, h->g, cv3<S>(xe), cv3<S>(xb),
                         cv3<S>(ue), cv3<S>(ub), cv3<S>(h->t0), h->has_t0, mv3<S>(ye), mv3<S>(yb), (float)cc,
                         (float)k2, (float)k3, af, 3 * h->g.cs);
    }
  }
  if (!h->comp)
    hipLaunchKernelGGL(k_rk3<PR>, box_grid(h), dim3(BX, BY), 0, h->stream, h->g, cv3<S>(xe), cv3<S>(xb),
                       cv3<S>(ue), cv3<S>(ub), cv3<S>(h->t0), h->has_t0, mv3<S>(ye), mv3<S>(yb), cc,
                       k1, k2, k3, ab);
  int rc = last_launch(h, "k_rk3");
  if (rc) return rc;
  if (h->timing) HIPCHK(h, hipEventRecord(e1, h->stream));
  return TFF_OK;
}

template <class PR>
int launch_clean(tff_handle* h, const Vec& xe, const Vec& xb, const Vec& ye) {
  using S = typename PR::S;
  PhaseScope phase(h, TFF_PHASE_SOLVER);
  if constexpr (std::is_same<S, float>::value) {
    if (h->comp) {
      hipLaunchKernelGGL(k_cleanc, box_grid(h), dim3(BX, BY), 0, h->stream, h->g, cv3<S>(xe), cv3<S>(xb),
                         mv3<S>(ye), h->cfg.absorb_fixed, 3 * h->g.cs);
      return last_launch(h, "k_cleanc");
    }
  }
  hipLaunchKernelGGL(k_clean<PR>, box_grid(h), dim3(BX, BY), 0, h->stream, h->g, cv3<S>(xe),
                     cv3<S>(xb), mv3<S>(ye), h->cfg.absorb_fixed);
  return last_launch(h, "k_clean");
}

template <class S>
int launch_fill(tff_handle* h, std::initializer_list<const Vec*> vecs) {
  const int* b = h->cfg.boundary;
  if (b[0] != TFF_PERIODIC && b[1] != TFF_PERIODIC && b[2] != TFF_PERIODIC) return TFF_OK;
  PhaseScope phase(h, TFF_PHASE_EXCHANGE);
  std::vector<S*> arr;
  for (const Vec* v : vecs)
    for (int c = 0; c < (h->comp ? 6 : 3); ++c) arr.push_back((S*)v->p[0] + (int64_t)c * h->g.cs);
  const int sx = h->g.nx + 2 * NG, sy = h->g.ny + 2 * NG;
  const int64_t nshell = (int64_t)2 * NG * sx * sy + (int64_t)2 * NG * sx * h->g.nz +
                         (int64_t)2 * NG * h->g.ny * h->g.nz;
  const int tb = 256;
  constexpr int CAP = (int)(sizeof(FillSet<S>::a) / sizeof(S*));
  for (size_t a0 = 0; a0 < arr.size(); a0 += CAP) {
    FillSet<S> fs{};
    fs.n = 0;
    for (size_t q = a0; q < arr.size() && fs.n < CAP; ++q) fs.a[fs.n++] = arr[q];
    hipLaunchKernelGGL(k_fill<S>, dim3((unsigned)((nshell + tb - 1) / tb)), dim3(tb), 0, h->stream,
                       h->g, b[0] == TFF_PERIODIC, b[1] == TFF_PERIODIC, b[2] == TFF_PERIODIC, fs,
                       nshell);
    int rc = last_launch(h, "k_fill");
    if (rc) return rc;
  }
  return TFF_OK;
}

template <class PR>
int launch_bc(tff_handle* h, const Vec& e, const Vec& b, const int* step_ptr, int step_val) {
  using S = typename PR::S;
  using P = typename PR::P;
  if (h->bc_kind == TFF_BC_NONE) return TFF_OK;
  PhaseScope phase(h, TFF_PHASE_HOOK);
  const int sx = h->g.nx + 2 * NG, sy = h->g.ny + 2 * NG;
  const int tb = 256;
  const P cc = std::is_same<S, P>::value ? (P)(S)h->cfg.cc : (P)h->cfg.cc;
  const int64_t lo = h->comp ? 3 * h->g.cs : 0;
  switch (h->bc_kind) {
    case TFF_BC_LOOP: {
      const int64_t n = (int64_t)(2 * NG + 1) * sx * sy;
      hipLaunchKernelGGL(k_bc_loop<PR>, dim3((unsigned)((n + tb - 1) / tb)), dim3(tb), 0, h->stream,
                         h->g, mv3<S>(e), mv3<S>(b), h->d_bc, cc, step_ptr, step_val, lo);
      return last_launch(h, "k_bc_loop");
    }
    case TFF_BC_UNIFORM: {
      const int64_t n = (int64_t)(2 * NG + 1) * sx * sy;
      hipLaunchKernelGGL(k_bc_uniform<PR>, dim3((unsigned)((n + tb - 1) / tb)), dim3(tb), 0,
                         h->stream, h->g, mv3<S>(e), mv3<S>(b), h->d_bc, cc, step_ptr, step_val, lo);
      return last_launch(h, "k_bc_uniform");
    }
    case TFF_BC_PSR: {
      const int64_t n = (int64_t)h->psr_n.x * h->psr_n.y * h->psr_n.z;
      if (n <= 0) return TFF_OK;
      const P ccinv = (P)1.0 / cc;
      hipLaunchKernelGGL(k_bc_psr_eval<PR>, dim3((unsigned)((n + tb - 1) / tb)), dim3(tb), 0,
                         h->stream, h->g, cv3<S>(e), cv3<S>(b), h->d_bc, ccinv, h->psr_lo, h->psr_n,
                         (S*)h->psr_stage, step_ptr, step_val);
      int rc = last_launch(h, "k_bc_psr_eval");
      if (rc) return rc;
      hipLaunchKernelGGL(k_bc_psr_commit<PR>, dim3((unsigned)((n + tb - 1) / tb)), dim3(tb), 0,
                         h->stream, h->g, mv3<S>(e), mv3<S>(b), h->d_bc, h->psr_lo, h->psr_n,
                         (const S*)h->psr_stage, lo);
      return last_launch(h, "k_bc_psr_commit");
    }
    default:
      return TFF_OK;
  }
}

template <class PR>
int ensure_t0(tff_handle* h) {
  using S = typename PR::S;
  if (!h->t0_dirty) return TFF_OK;
  if (!h->t0.p[0]) {
    int rc = alloc_vec(h, h->t0);
    if (rc) return rc;
  }
  HIPCHK(h, hipMemsetAsync(h->d_flag, 0, sizeof(int), h->stream));
  dim3 grid((h->g.nx + BX - 1) / BX, (h->g.ny + BY - 1) / BY, h->g.nz);
  hipLaunchKernelGGL(k_b0_curl<S>, grid, dim3(BX, BY), 0, h->stream, h->g, cv3<S>(h->b0),
                     mv3<S>(h->t0), h->d_flag);
  int rc = last_launch(h, "k_b0_curl");
  if (rc) return rc;
  int flag = 0;
  HIPCHK(h, hipMemcpyAsync(&flag, h->d_flag, sizeof(int), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  if (flag != h->has_t0) invalidate_graphs(h);
  h->has_t0 = flag;
  h->t0_dirty = false;
  return TFF_OK;
}

// One timestep in the order of tristanmainloop.F90:50-83.  X denotes the
// cleaned fields a stage reads, Y the raw stage output.
template <class PR>
int enqueue_step(tff_handle* h, const int* step_ptr, int step_val) {
  using S = typename PR::S;
  const double* cf = nullptr;
  static const double c64[9] = {1.0, 0.0, 1.0, 3.0 / 4.0, 1.0 / 4.0, 1.0 / 4.0,
                                1.0 / 3.0, 2.0 / 3.0, 2.0 / 3.0};
  cf = c64;  // rounded to S inside launch_rk3: FP32 gets 1.0/3.0 as float
  int rc;
  // stage 1: Y1 = Un + L(Un)
  if ((rc = launch_rk3<PR>(h, h->es, h->bs, h->es, h->bs, h->w1e, h->w1b, cf[0], cf[1], cf[2]))) return rc;
  if ((rc = launch_bc<PR>(h, h->w1e, h->w1b, step_ptr, step_val))) return rc;
  if ((rc = launch_fill<S>(h, {&h->w1e, &h->w1b}))) return rc;
  if ((rc = launch_clean<PR>(h, h->w1e, h->w1b, h->ce))) return rc;
  if ((rc = launch_fill<S>(h, {&h->ce}))) return rc;
  // stage 2: Y2 = 3/4 Un + 1/4 X1 + 1/4 L(X1)
  if ((rc = launch_rk3<PR>(h, h->ce, h->w1b, h->es, h->bs, h->w2e, h->w2b, cf[3], cf[4], cf[5]))) return rc;
  if ((rc = launch_bc<PR>(h, h->w2e, h->w2b, step_ptr, step_val))) return rc;
  if ((rc = launch_fill<S>(h, {&h->w2e, &h->w2b}))) return rc;
  if ((rc = launch_clean<PR>(h, h->w2e, h->w2b, h->ce))) return rc;
  if ((rc = launch_fill<S>(h, {&h->ce}))) return rc;
  // stage 3: Y3 = 1/3 Un + 2/3 X2 + 2/3 L(X2), written over Un in place
  // (each thread reads Un only at its own cell before writing it)
  if ((rc = launch_rk3<PR>(h, h->ce, h->w2b, h->es, h->bs, h->es, h->bs, cf[6], cf[7], cf[8]))) return rc;
  if ((rc = launch_bc<PR>(h, h->es, h->bs, step_ptr, step_val))) return rc;
  if ((rc = launch_fill<S>(h, {&h->es, &h->bs}))) return rc;
  if ((rc = launch_clean<PR>(h, h->es, h->bs, h->ce))) return rc;
  if ((rc = launch_fill<S>(h, {&h->ce}))) return rc;
  return TFF_OK;
}

// ---- precision dispatch ----------------------------------------------------
template <template <class> class F, class... A>
int dispatch(tff_handle* h, A... a) {
  switch (h->cfg.precision) {
    case TFF_FP64: return F<PrecF64>::run(h, a...);
    case TFF_FP32: return F<PrecF32>::run(h, a...);
    default: return F<PrecF32Bulk>::run(h, a...);
  }
}

// ---- fused path -------------------------------------------------------------
// Tile shapes the fused stage kernel is instantiated for; the autotuner
// picks one (and the z-chunk) per handle.
// Taller tiles clean a smaller halo ring per cell (32x16: 0.20 ring cells
// per cell, 64x8: 0.29); 1024-thread tiles run one workgroup per CU.
using TileA = Tile<32, 4>;
using TileB = Tile<32, 8>;
using TileC = Tile<64, 4>;
using TileD = Tile<64, 8>;
using TileE = Tile<32, 16>;
using TileF = Tile<64, 16>;
constexpr int NTILES = 6;
// steps per captured graph in tff_run (fewer graph launches and step-counter
// updates; a remainder replays the one-step graph)
constexpr int RUN_GRAPH_STEPS = 8;
constexpr int TILE_TX[NTILES] = {32, 32, 64, 64, 32, 64};
constexpr int TILE_TY[NTILES] = {4, 8, 4, 8, 16, 16};

// f(TL{}) for tile shape t
template <class R, class F>
R with_tile(int t, F&& f) {
  switch (t) {
    case 0: return f(TileA{});
    case 1: return f(TileB{});
    case 2: return f(TileC{});
    case 3: return f(TileD{});
    case 4: return f(TileE{});
    default: return f(TileF{});
  }
}

// whether a TL workgroup's LDS fits a CU at all (shapes that do not are
// never instantiated for that precision)
constexpr size_t CU_LDS = 160 * 1024;
template <class PR, class TL>
constexpr bool tile_fits() {
  return stage_lds_max<PR, TL>() <= CU_LDS;
}

template <class PR>
size_t tile_lds(int t) {
  return with_tile<size_t>(t, [](auto tl) { return stage_lds_max<PR, decltype(tl)>(); });
}

// launch kernel `k` (a stage-kernel instance) with `lds` bytes of dynamic
// LDS, raising its LDS limit on the first launch
template <class K, class A>
void launch_lds(K k, bool& attr, unsigned grid, unsigned block, size_t lds, hipStream_t st, const A& a) {
  if (!attr) {
    hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    attr = true;
  }
  hipLaunchKernelGGL(k, dim3(grid), dim3(block), lds, st, a);
}

bool all_periodic(const tff_handle* h) {
  return h->cfg.boundary[0] == TFF_PERIODIC && h->cfg.boundary[1] == TFF_PERIODIC &&
         h->cfg.boundary[2] == TFF_PERIODIC;
}
// every ghost read is a periodic wrap and no hook writes ghosts
bool wrap_mode(const tff_handle* h) { return all_periodic(h) && h->bc_kind == TFF_BC_NONE; }

// A hook that reads no field value (none, the uniform twist, the driven
// loop): it cannot see a carried value that the periodic fill overwrites.
bool blind_hook(const tff_handle* h) {
  return h->bc_kind == TFF_BC_NONE || h->bc_kind == TFF_BC_UNIFORM ||
         (h->bc_kind == TFF_BC_LOOP && h->bc_host[8] == 2.0);
}

// x and y periodic, z open, a blind hook: the fused path reads the x/y
// ghosts of the interior planes as periodic wraps and stores only the
// z-ghost planes (WRAP_MIXED)
bool mixed_mode(const tff_handle* h) {
  const int* b = h->cfg.boundary;
  return b[0] == TFF_PERIODIC && b[1] == TFF_PERIODIC && b[2] == TFF_OPEN && blind_hook(h) && !h->comp;
}

// the ghost mode of the fused stage kernel (WRAP_STORED, _PERIODIC, _MIXED)
int ghost_mode(const tff_handle* h) {
  return wrap_mode(h) ? WRAP_PERIODIC : (mixed_mode(h) ? WRAP_MIXED : WRAP_STORED);
}

int64_t shell_count(const tff_handle* h) {
  const int sx = h->g.nx + 2 * NG, sy = h->g.ny + 2 * NG;
  return (int64_t)2 * NG * sx * sy + (int64_t)2 * NG * sx * h->g.nz + (int64_t)2 * NG * h->g.ny * h->g.nz;
}

// the tile grid of a TL launch: x fastest, then y, then z-chunks
template <class TL, class SA>
SA with_tile_grid(const tff_handle* h, const SA& a) {
  SA b = a;
  b.nbx = (h->g.nx + TL::TX - 1) / TL::TX;
  b.nby = (h->g.ny + TL::TY - 1) / TL::TY;
  b.nblocks = b.nbx * b.nby * ((h->g.nz + a.kchunk - 1) / a.kchunk);
  return b;
}

template <class PR, class TL>
int launch_stage_t(tff_handle* h, const StageArgs<typename PR::S, typename PR::P>& a) {
  if constexpr (!tile_fits<PR, TL>()) return fail(h, TFF_EINVAL, "tile shape exceeds the LDS of a CU");
  else {
  const size_t lds = stage_lds<PR, TL, PR::null_fp64>();
  static bool attr_p = false, attr_s = false;
  const auto b = with_tile_grid<TL>(h, a);
  if (PR::null_fp64 && (!b.fixmask || b.nblocks > h->fix_cap))
    return fail(h, TFF_ESTATE, "near-null plane masks missing or too small");
  const unsigned grid = (unsigned)((b.nblocks + 7) / 8 * 8);
  static bool attr_m = false;
  if (b.wrap == WRAP_PERIODIC)
    launch_lds(k_stage<PR, TL, WRAP_PERIODIC>, attr_p, grid, TL::NT, lds, h->stream, b);
  else if (b.wrap == WRAP_MIXED)
    launch_lds(k_stage<PR, TL, WRAP_MIXED>, attr_m, grid, TL::NT, lds, h->stream, b);
  else
    launch_lds(k_stage<PR, TL, WRAP_STORED>, attr_s, grid, TL::NT, lds, h->stream, b);
  return last_launch(h, "k_stage");
  }
}

// k_stage_fix after a k_stage<PR, TL> launch: up to one block per CU scans
// the tiles' plane masks (a few words each) and recomputes what they mark.
constexpr unsigned FIX_GRID = 256;
template <class PR, class TL>
int launch_fix_t(tff_handle* h, const StageArgs<typename PR::S, typename PR::P>& a) {
  const size_t lds = stage_lds<PR, TL, false>();
  static bool attr_p = false, attr_s = false;
  const auto b = with_tile_grid<TL>(h, a);
  const unsigned grid = (unsigned)std::min<int64_t>(FIX_GRID, b.nblocks);
  static bool attr_m = false;
  if (b.wrap == WRAP_PERIODIC)
    launch_lds(k_stage_fix<PR, TL, WRAP_PERIODIC>, attr_p, grid, TL::NT, lds, h->stream, b);
  else if (b.wrap == WRAP_MIXED)
    launch_lds(k_stage_fix<PR, TL, WRAP_MIXED>, attr_m, grid, TL::NT, lds, h->stream, b);
  else
    launch_lds(k_stage_fix<PR, TL, WRAP_STORED>, attr_s, grid, TL::NT, lds, h->stream, b);
  return last_launch(h, "k_stage_fix");
}

template <class PR>
int launch_stage(tff_handle* h, const StageArgs<typename PR::S, typename PR::P>& a, int tile) {
  PhaseScope phase(h, TFF_PHASE_SOLVER);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (h->timing) {
    while (h->tpool.size() < h->tused + 2) {
      hipEvent_t e;
      // timing only: no system-scope fence (its cache writeback and
      // invalidation would slow the kernel it brackets)
      HIPCHK(h, hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
      h->tpool.push_back(e);
    }
    e0 = h->tpool[h->tused++];
    e1 = h->tpool[h->tused++];
    HIPCHK(h, hipEventRecord(e0, h->stream));
  }
  int rc = with_tile<int>(tile, [&](auto tl) { return launch_stage_t<PR, decltype(tl)>(h, a); });
  if (rc) return rc;
  if (h->timing) HIPCHK(h, hipEventRecord(e1, h->stream));
  if constexpr (PR::null_fp64)
    rc = with_tile<int>(tile, [&](auto tl) { return launch_fix_t<PR, decltype(tl)>(h, a); });
  return rc;
}

template <class PR>
StageArgs<typename PR::S, typename PR::P> stage_common(tff_handle* h) {
  using S = typename PR::S;
  using P = typename PR::P;
  StageArgs<S, P> a{};
  a.g = h->g;
  const bool exact = std::is_same<S, P>::value;
  a.cc = exact ? (P)(S)h->cfg.cc : (P)h->cfg.cc;
  a.ab.on = h->cfg.boundary[0] == TFF_ABSORB;
  a.ab.fixed = h->cfg.absorb_fixed;
  a.ab.xc = (P)0.5 * (P)h->g.nx;
  a.ab.yc = (P)0.5 * (P)h->g.ny;
  a.ab.zc = (P)0.5 * (P)h->g.nz;
  a.ab.rabsorb = std::min(std::min(a.ab.xc, a.ab.yc), a.ab.zc) - (P)h->cfg.absorb_buff;
  a.ab.buff = (P)h->cfg.absorb_buff;
  a.t0 = base_c<S>(h->t0);
  a.has_t0 = h->has_t0;
  a.guard = h->cfg.absorb_fixed;
  a.wrap = ghost_mode(h);
  a.kchunk = h->kchunk;
  a.fixmask = h->d_fixmask;
  return a;
}

template <class PR>
void set_coef(tff_handle* h, StageArgs<typename PR::S, typename PR::P>& a, double c1, double c2, double c3) {
  using S = typename PR::S;
  using P = typename PR::P;
  const bool exact = std::is_same<S, P>::value;
  a.c1 = exact ? (P)(S)c1 : (P)c1;
  a.c2 = exact ? (P)(S)c2 : (P)c2;
  a.c3 = exact ? (P)(S)c3 : (P)c3;
}

// Whether a box's ghost cells keep their values through a run of fused
// steps: with no periodic axis nothing fills them, rk3Update and cleanEpar
// leave them alone (a non-filled ghost's cleaned value is its raw value),
// and the only hook that may run writes interior cells only
// (user_psr.F90:72-332).  Once one full step has carried the ghosts
// through every stage buffer, carry and ghost clean rewrite what is there.
bool ghosts_static(const tff_handle* h) {
  const int* b = h->cfg.boundary;
  return b[0] != TFF_PERIODIC && b[1] != TFF_PERIODIC && b[2] != TFF_PERIODIC &&
         (h->bc_kind == TFF_BC_NONE || h->bc_kind == TFF_BC_PSR);
}

// After a fused stage wrote the interior of `oe/ob`: the reference's hook,
// ghost fill and (for the next stage) the cleaned E of the ghost cells.
template <class PR>
int ghost_passes(tff_handle* h, const Vec& ge_in, const Vec& prev_b, const Vec& oe, const Vec& ob,
                 const int* step_ptr, int step_val) {
  using S = typename PR::S;
  if (h->ghosts_settled) return launch_bc<PR>(h, oe, ob, step_ptr, step_val);
  const int64_t n = shell_count(h);
  const int tb = 256;
  const unsigned nb = (unsigned)((n + tb - 1) / tb);
  // a blind hook cannot see a carried value that the periodic fill then
  // overwrites
  const int* bd = h->cfg.boundary;
  const int3 per = blind_hook(h) ? int3{bd[0] == TFF_PERIODIC, bd[1] == TFF_PERIODIC, bd[2] == TFF_PERIODIC}
                                 : int3{0, 0, 0};
  int rc;
  {
    PhaseScope phase(h, TFF_PHASE_EXCHANGE);
    hipLaunchKernelGGL(k_carry<S>, dim3(nb), dim3(tb), 0, h->stream, h->g, cv3<S>(ge_in), cv3<S>(prev_b),
                       mv3<S>(oe), mv3<S>(ob), n, per);
    rc = last_launch(h, "k_carry");
  }
  if (rc) return rc;
  if ((rc = launch_bc<PR>(h, oe, ob, step_ptr, step_val))) return rc;
  if ((rc = launch_fill<S>(h, {&oe, &ob}))) return rc;
  const int* b = h->cfg.boundary;
  PhaseScope phase(h, TFF_PHASE_EXCHANGE);
  hipLaunchKernelGGL(k_ghost_clean<PR>, dim3(nb), dim3(tb), 0, h->stream, h->g, b[0] == TFF_PERIODIC,
                     b[1] == TFF_PERIODIC, b[2] == TFF_PERIODIC, cv3<S>(oe), cv3<S>(ob), mv3<S>(h->ge),
                     h->cfg.absorb_fixed, n);
  return last_launch(h, "k_ghost_clean");
}

// The ghost passes of a WRAP_MIXED box after a fused stage: only the
// z-ghost planes are stored.  They start as the stage input's values there
// (rk3Update leaves ghosts alone; nothing cleans or fills those planes, so
// the input's raw E is its cleaned E there), then the hook writes them, and
// k_edge_clean stores the cleaned E of the x/y ghost ring of planes 0 and
// nz-1 for the next stage.  No fill: the stage kernel reads the interior
// planes' x/y ghosts as wraps, and exchangeFields leaves the edges of the
// open z axis alone.
template <class PR>
int ghost_passes_mixed(tff_handle* h, const Vec& ye_in, const Vec& prev_b, const Vec& oe, const Vec& ob,
                       const int* step_ptr, int step_val) {
  using S = typename PR::S;
  // the z-ghost planes come first in shell_cell's order
  const int64_t nzc = (int64_t)2 * NG * (h->g.nx + 2 * NG) * (h->g.ny + 2 * NG);
  const int tb = 256;
  int rc;
  {
    PhaseScope phase(h, TFF_PHASE_EXCHANGE);
    hipLaunchKernelGGL(k_carry<S>, dim3((unsigned)((nzc + tb - 1) / tb)), dim3(tb), 0, h->stream, h->g,
                       cv3<S>(ye_in), cv3<S>(prev_b), mv3<S>(oe), mv3<S>(ob), nzc, int3{1, 1, 0});
    rc = last_launch(h, "k_carry");
  }
  if (rc) return rc;
  if ((rc = launch_bc<PR>(h, oe, ob, step_ptr, step_val))) return rc;
  PhaseScope phase(h, TFF_PHASE_EXCHANGE);
  const int ne = (2 * (h->g.nx + 2) + 2 * h->g.ny) * (h->g.nz > 1 ? 2 : 1);
  hipLaunchKernelGGL(k_edge_clean<PR>, dim3((unsigned)((ne + tb - 1) / tb)), dim3(tb), 0, h->stream, h->g,
                     base_c<S>(oe), base_c<S>(ob), base_m<S>(h->ge), h->cfg.absorb_fixed);
  return last_launch(h, "k_edge_clean");
}

// One timestep on the fused path.  Buffers: es/bs hold the state (raw
// output of the last stage unless state_clean), w1/w2 the stage outputs,
// ce the cleaned U^n E, ge the cleaned ghost E.
template <class PR>
int enqueue_fused_step(tff_handle* h, const int* step_ptr, int step_val) {
  using S = typename PR::S;
  if (!h->ge.p[0] || h->tile < 0 || h->tile >= NTILES || h->kchunk < 1)
    return fail(h, TFF_ESTATE, "fused path not set up (ghost buffer or tuning missing)");
  const bool wrap = wrap_mode(h), mixed = mixed_mode(h);
  const bool first = h->state_clean;
  int rc;
  auto a = stage_common<PR>(h);
  // the ghost passes after a stage that read (ye, yb) and wrote (oe, ob);
  // ge_in: the cleaned E of the input's ghosts (stored mode)
  auto ghosts = [&](const Vec& ye, const Vec& ge_in, const Vec& yb, const Vec& oe, const Vec& ob) {
    if (wrap) return (int)TFF_OK;
    if (mixed) return ghost_passes_mixed<PR>(h, ye, yb, oe, ob, step_ptr, step_val);
    return ghost_passes<PR>(h, ge_in, yb, oe, ob, step_ptr, step_val);
  };
  // stage 1: U1 = Un + L(Un), Un = clean(state) (or the state itself when it
  // is already clean)
  set_coef<PR>(h, a, 1.0, 0.0, 1.0);
  a.ye = base_c<S>(h->es);
  a.yb = base_c<S>(h->bs);
  a.ge = first ? base_c<S>(h->es) : base_c<S>(h->ge);
  a.stage1 = 1;
  a.clean_in = first ? 0 : 1;
  a.oe = base_m<S>(h->w1e);
  a.ob = base_m<S>(h->w1b);
  a.u0e = first ? nullptr : base_m<S>(h->ce);
  if ((rc = launch_stage<PR>(h, a, h->tile))) return rc;
  const Vec& u0e = first ? h->es : h->ce;
  if ((rc = ghosts(h->es, first ? h->es : h->ge, h->bs, h->w1e, h->w1b))) return rc;
  // stage 2: U2 = 3/4 Un + 1/4 X1 + 1/4 L(X1)
  set_coef<PR>(h, a, 3.0 / 4.0, 1.0 / 4.0, 1.0 / 4.0);
  a.ye = base_c<S>(h->w1e);
  a.yb = base_c<S>(h->w1b);
  a.ge = base_c<S>(h->ge);
  a.stage1 = 0;
  a.clean_in = 1;
  a.ue = base_c<S>(u0e);
  a.ub = base_c<S>(h->bs);
  a.oe = base_m<S>(h->w2e);
  a.ob = base_m<S>(h->w2b);
  a.u0e = nullptr;
  if ((rc = launch_stage<PR>(h, a, h->tile))) return rc;
  if ((rc = ghosts(h->w1e, h->ge, h->w1b, h->w2e, h->w2b))) return rc;
  // stage 3: U3 = 1/3 Un + 2/3 X2 + 2/3 L(X2), written over the state in
  // place (each thread reads Un only at its own cell before writing it)
  set_coef<PR>(h, a, 1.0 / 3.0, 2.0 / 3.0, 2.0 / 3.0);
  a.ye = base_c<S>(h->w2e);
  a.yb = base_c<S>(h->w2b);
  a.oe = base_m<S>(h->es);
  a.ob = base_m<S>(h->bs);
  if ((rc = launch_stage<PR>(h, a, h->tile))) return rc;
  if ((rc = ghosts(h->w2e, h->ge, h->w2b, h->es, h->bs))) return rc;
  h->state_clean = false;
  return TFF_OK;
}

// Turn the fused path's raw state back into the reference's arrays: the
// cleaned interior, cleaned ghosts (periodic wrap, or the ghost-pass
// values) and filled B ghosts.
template <class PR>
int ensure_clean_t(tff_handle* h) {
  using S = typename PR::S;
  if (h->state_clean) return TFF_OK;
  int rc;
  // in wrap mode the raw state's ghosts were never written, in mixed mode
  // its x/y ghosts of the interior planes: fill them first (the reference's
  // fill before cleanEpar), then clean and fill again
  const bool wrap = wrap_mode(h), mixed = mixed_mode(h);
  if ((wrap || mixed) && (rc = launch_fill<S>(h, {&h->es, &h->bs}))) return rc;
  if ((rc = launch_clean<PR>(h, h->es, h->bs, h->ce))) return rc;
  if (wrap) {
    if ((rc = launch_fill<S>(h, {&h->ce}))) return rc;
  } else if (mixed) {
    // the z-ghost planes (edges included) are never cleaned: every ghost
    // takes the raw value, then the fill puts the cleaned wraps on the x/y
    // ghosts of the interior planes
    const int64_t n = shell_count(h);
    const int tb = 256;
    hipLaunchKernelGGL(k_carry<S>, dim3((unsigned)((n + tb - 1) / tb)), dim3(tb), 0, h->stream, h->g,
                       cv3<S>(h->es), cv3<S>(h->bs), mv3<S>(h->ce), V3<S>{nullptr, nullptr, nullptr}, n,
                       int3{0, 0, 0});
    if ((rc = last_launch(h, "k_carry"))) return rc;
    if ((rc = launch_fill<S>(h, {&h->ce}))) return rc;
  } else {
    const int64_t n = shell_count(h);
    const int tb = 256;
    hipLaunchKernelGGL(k_carry<S>, dim3((unsigned)((n + tb - 1) / tb)), dim3(tb), 0, h->stream, h->g,
                       cv3<S>(h->ge), cv3<S>(h->bs), mv3<S>(h->ce), V3<S>{nullptr, nullptr, nullptr}, n,
                       int3{0, 0, 0});
    if ((rc = last_launch(h, "k_carry"))) return rc;
  }
  std::swap(h->es, h->ce);
  h->state_clean = true;
  invalidate_graphs(h);
  return TFF_OK;
}

template <class PR>
struct DoEnsureClean {
  static int run(tff_handle* h) { return ensure_clean_t<PR>(h); }
};

// Autotuner: time every tile shape that fits the LDS limit with every
// z-chunk on one stage-2 launch (state buffers read, work buffers written,
// nothing the state depends on), keep the fastest.
template <class PR>
int autotune_t(tff_handle* h) {
  using S = typename PR::S;
  if (h->tile >= 0) return TFF_OK;
  // z-chunks of 4 planes spread a small box (up to 128^3) over more
  // workgroups: a tile's planes are a chain of dependent iterations, so on a
  // box that does not fill the CUs the chain length sets the launch time
  const int kcs[] = {4, 8, 16, 32, 64};
  const bool small = (int64_t)h->g.nx * h->g.ny * h->g.nz <= (int64_t)128 * 128 * 128;
  float best = 1e30f;
  int bt = -1, bk = 0;
  hipEvent_t e0, e1;
  HIPCHK(h, hipEventCreate(&e0));
  HIPCHK(h, hipEventCreate(&e1));
  // no per-launch or per-phase events while the search launches (nor inside
  // its captured graphs); the search's time is not the steps' time
  const bool saved_timing = h->timing, saved_phases = h->phases;
  h->timing = false;
  h->phases = false;
  // every (tile, z-chunk) the box and the device admit
  std::vector<std::pair<int, int>> cand;
  for (int t = 0; t < NTILES; ++t) {
    if (tile_lds<PR>(t) > (size_t)h->max_lds) continue;
    for (int kc : kcs) {
      if ((kc > 2 * h->g.nz && kc != 8) || (kc == 4 && !small)) continue;
      cand.emplace_back(t, kc);
    }
  }
  // Two passes over the candidates, each timing one replay of a graph of
  // three launches of every shape (captured after one eager warm-up launch
  // in the first pass); a shape keeps its faster pass.  Replayed graphs
  // rank the shapes as the graph-replayed steps run them (eager launches
  // favoured shorter z-chunks), the faster of two passes keeps shapes
  // within a few per cent of each other from swapping places run to run,
  // and timing the passes one after the other keeps a drift of the GPU's
  // clocks during the search from favouring the shapes timed last.
  std::vector<float> ms(cand.size(), 1e30f);
  std::vector<hipGraphExec_t> execs(cand.size(), nullptr);
  auto drop_execs = [&]() {
    for (auto& x : execs)
      if (x) hipGraphExecDestroy(x);
  };
  int rc = 0;
  for (int pass = 0; pass < 2 && rc == 0; ++pass) {
    for (size_t c = 0; c < cand.size() && rc == 0; ++c) {
      const int t = cand[c].first;
      h->kchunk = cand[c].second;
      if (pass == 0) {
        auto a = stage_common<PR>(h);
        set_coef<PR>(h, a, 3.0 / 4.0, 1.0 / 4.0, 1.0 / 4.0);
        a.ye = base_c<S>(h->es);
        a.yb = base_c<S>(h->bs);
        a.ge = base_c<S>(h->es);
        a.clean_in = 1;
        a.ue = base_c<S>(h->es);
        a.ub = base_c<S>(h->bs);
        a.oe = base_m<S>(h->w2e);
        a.ob = base_m<S>(h->w2b);
        if ((rc = launch_stage<PR>(h, a, t))) break;  // warm-up (instruction cache, kernel attributes)
        hipGraph_t graph = nullptr;
        if (hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) {
          rc = fail(h, TFF_EHIP, "autotune: hipStreamBeginCapture failed");
          break;
        }
        for (int r = 0; r < 3 && rc == 0; ++r) rc = launch_stage<PR>(h, a, t);
        const hipError_t e = hipStreamEndCapture(h->stream, &graph);
        if (rc == 0 && e != hipSuccess) rc = fail(h, TFF_EHIP, "autotune: hipStreamEndCapture failed");
        if (rc == 0 && hipGraphInstantiate(&execs[c], graph, nullptr, nullptr, 0) != hipSuccess)
          rc = fail(h, TFF_EHIP, "autotune: hipGraphInstantiate failed");
        if (graph) hipGraphDestroy(graph);
        if (rc) break;
      }
      float m = 0.f;
      if (hipEventRecord(e0, h->stream) != hipSuccess || hipGraphLaunch(execs[c], h->stream) != hipSuccess ||
          hipEventRecord(e1, h->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
          hipEventElapsedTime(&m, e0, e1) != hipSuccess) {
        rc = fail(h, TFF_EHIP, "autotune: timed replay failed");
        break;
      }
      ms[c] = std::min(ms[c], m);
    }
  }
  drop_execs();
  h->timing = saved_timing;
  h->phases = saved_phases;
  if (rc) {
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    return rc;
  }
  for (size_t c = 0; c < cand.size(); ++c) {
    if (ms[c] < best) {
      best = ms[c];
      bt = cand[c].first;
      bk = cand[c].second;
    }
  }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  if (bt < 0) return fail(h, TFF_EINVAL, "no tile shape fits the device's LDS");
  h->tile = bt;
  h->kchunk = bk;
  return TFF_OK;
}

// One timestep, or with use_graph up to `count` steps from one captured
// graph; *done gets the number of steps enqueued.  Under a graph the hooks
// read the step from d_step plus their step's offset in the graph, and one
// k_bump_step at the end advances d_step by the steps captured.
template <class PR>
struct DoStep {
  static int run(tff_handle* h, int step, int use_graph, int count, int* done) {
    *done = 1;
    int rc = ensure_t0<PR>(h);
    if (rc) return rc;
    if (h->fused) {
      if ((rc = autotune_t<PR>(h))) return rc;
      // the first step after the state was (re)made clean runs eagerly; the
      // steady state has fixed buffer roles and is replayed from one graph
      if (!use_graph || h->timing || h->phases || h->state_clean) {
        if ((rc = enqueue_fused_step<PR>(h, nullptr, step))) return rc;
        // this step carried the ghosts through w1, w2 and the state
        if (ghosts_static(h)) h->ghosts_settled = true;
        if (use_graph) hipLaunchKernelGGL(k_bump_step, dim3(1), dim3(64), 0, h->stream, h->d_step, 1);
        return last_launch(h, "k_bump_step");
      }
    } else {
      if ((rc = ensure_clean_t<PR>(h))) return rc;
      if (!use_graph || h->timing || h->phases) {
        rc = enqueue_step<PR>(h, nullptr, step);
        if (rc) return rc;
        std::swap(h->es, h->ce);
        if (use_graph) hipLaunchKernelGGL(k_bump_step, dim3(1), dim3(64), 0, h->stream, h->d_step, 1);
        return last_launch(h, "k_bump_step");
      }
      count = 1;  // the buffer roles alternate from step to step
    }
    auto key = std::make_tuple(h->es.p[0], h->ce.p[0], count);
    auto it = h->graphs.find(key);
    if (it == h->graphs.end()) {
      if (!h->retired.empty()) {
        HIPCHK(h, hipStreamSynchronize(h->stream));
        drop_retired(h);
      }
      hipGraph_t graph = nullptr;
      HIPCHK(h, hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
      for (int s = 0; s < count && rc == TFF_OK; ++s)
        rc = h->fused ? enqueue_fused_step<PR>(h, h->d_step, s) : enqueue_step<PR>(h, h->d_step, s);
      if (rc == TFF_OK) hipLaunchKernelGGL(k_bump_step, dim3(1), dim3(64), 0, h->stream, h->d_step, count);
      hipError_t e = hipStreamEndCapture(h->stream, &graph);
      if (rc) {
        // keep the enqueue's own error; the partial graph is not used
        if (e == hipSuccess && graph) hipGraphDestroy(graph);
        return rc;
      }
      if (e != hipSuccess) return fail(h, TFF_EHIP, "hipStreamEndCapture failed");
      hipGraphExec_t exec;
      e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
      hipGraphDestroy(graph);
      if (e != hipSuccess) return fail(h, TFF_EHIP, "hipGraphInstantiate failed");
      it = h->graphs.emplace(key, exec).first;
    }
    HIPCHK(h, hipGraphLaunch(it->second, h->stream));
    if (!h->fused) std::swap(h->es, h->ce);
    *done = count;
    return TFF_OK;
  }
};

// The output datasets of writeFields_hdf5 (output_istep >= 2): the cleaned
// state's node interpolants at every istep-th node, into host buffers.
int output_dims(const tff_handle* h, int istep, int d[3]) {
  d[0] = (h->g.nx + istep - 1) / istep;
  d[1] = (h->g.ny + istep - 1) / istep;
  d[2] = (h->g.nz + istep - 1) / istep;
  return TFF_OK;
}

// The raw components and rho at the output nodes (k_output_samples) into
// the caller's seven host buffers, staged through `dbuf`.
template <class PR>
struct DoSamples {
  static int run(tff_handle* h, int istep, void* const* out, int host_dtype, void* dbuf) {
    using S = typename PR::S;
    int d[3];
    output_dims(h, istep, d);
    const int64_t nout = (int64_t)d[0] * d[1] * d[2];
    const size_t hsz = host_dtype == TFF_HOST_F64 ? 8 : 4;
    const dim3 grid((unsigned)((d[0] + 63) / 64), (unsigned)((d[1] + 3) / 4), (unsigned)d[2]);
    if (hsz == 8)
      hipLaunchKernelGGL((k_output_samples<S, double>), grid, dim3(256), 0, h->stream, h->g, istep, d[0], d[1],
                         d[2], cv3<S>(h->es), cv3<S>(h->bs), (double*)dbuf);
    else
      hipLaunchKernelGGL((k_output_samples<S, float>), grid, dim3(256), 0, h->stream, h->g, istep, d[0], d[1],
                         d[2], cv3<S>(h->es), cv3<S>(h->bs), (float*)dbuf);
    int rc = last_launch(h, "k_output_samples");
    for (int f = 0; f < 7 && !rc; ++f)
      if (hipMemcpyAsync(out[f], (char*)dbuf + (size_t)f * nout * hsz, (size_t)nout * hsz, hipMemcpyDeviceToHost,
                         h->stream) != hipSuccess)
        rc = fail(h, TFF_EHIP, "hipMemcpyAsync of an output sample set failed");
    if (hipStreamSynchronize(h->stream) != hipSuccess && !rc) rc = fail(h, TFF_EHIP, "output: stream failed");
    return rc;
  }
};

// Bytes of device staging the six output datasets of one call take.
size_t output_bytes(const tff_handle* h, int istep, int host_dtype) {
  int d[3];
  output_dims(h, istep, d);
  return 6 * (size_t)d[0] * d[1] * d[2] * (host_dtype == TFF_HOST_F64 ? 8 : 4);
}

// The six datasets into the device staging `dbuf` (k_output), then into the
// caller's host buffers; returns when they are there.
template <class PR>
struct DoOutput {
  static int run(tff_handle* h, int istep, void* const* out, int host_dtype, void* dbuf) {
    using S = typename PR::S;
    int d[3];
    output_dims(h, istep, d);
    const int64_t nout = (int64_t)d[0] * d[1] * d[2];
    const size_t hsz = host_dtype == TFF_HOST_F64 ? 8 : 4;
    const dim3 grid((unsigned)((d[0] + 63) / 64), (unsigned)((d[1] + 3) / 4), (unsigned)d[2]);
    if (hsz == 8)
      hipLaunchKernelGGL((k_output<S, double>), grid, dim3(256), 0, h->stream, h->g, istep, d[0], d[1], d[2],
                         cv3<S>(h->es), cv3<S>(h->bs), (double*)dbuf);
    else
      hipLaunchKernelGGL((k_output<S, float>), grid, dim3(256), 0, h->stream, h->g, istep, d[0], d[1], d[2],
                         cv3<S>(h->es), cv3<S>(h->bs), (float*)dbuf);
    int rc = last_launch(h, "k_output");
    for (int f = 0; f < 6 && !rc; ++f)
      if (hipMemcpyAsync(out[f], (char*)dbuf + (size_t)f * nout * hsz, (size_t)nout * hsz, hipMemcpyDeviceToHost,
                         h->stream) != hipSuccess)
        rc = fail(h, TFF_EHIP, "hipMemcpyAsync of an output dataset failed");
    if (hipStreamSynchronize(h->stream) != hipSuccess && !rc) rc = fail(h, TFF_EHIP, "output: stream failed");
    return rc;
  }
};

template <class PR>
struct DoRk3 {
  static int run(tff_handle* h, double c1, double c2, double c3) {
    int rc = ensure_t0<PR>(h);
    if (rc) return rc;
    rc = launch_rk3<PR>(h, h->es, h->bs, h->en, h->bn, h->w1e, h->w1b, c1, c2, c3);
    if (rc) return rc;
    std::swap(h->es, h->w1e);
    std::swap(h->bs, h->w1b);
    invalidate_graphs(h);
    return TFF_OK;
  }
};

template <class PR>
struct DoClean {
  static int run(tff_handle* h) {
    int rc = launch_clean<PR>(h, h->es, h->bs, h->ce);
    if (rc) return rc;
    std::swap(h->es, h->ce);
    return TFF_OK;
  }
};

template <class PR>
struct DoBC {
  static int run(tff_handle* h, int step) { return launch_bc<PR>(h, h->es, h->bs, nullptr, step); }
};

template <class PR>
struct DoFill {
  static int run(tff_handle* h, int b0) {
    using S = typename PR::S;
    if (b0) {
      h->t0_dirty = true;
      return launch_fill<S>(h, {&h->es, &h->bs, &h->b0});
    }
    return launch_fill<S>(h, {&h->es, &h->bs});
  }
};

template <class PR>
struct DoReduce {
  static int run(tff_handle* h, double* out) {
    using S = typename PR::S;
    hipLaunchKernelGGL(k_reduce_partial<S>, dim3(h->red_blocks), dim3(RED_THREADS), 0, h->stream,
                       h->g, cv3<S>(h->es), cv3<S>(h->bs), h->d_part, h->comp ? 3 * h->g.cs : (int64_t)0);
    int rc = last_launch(h, "k_reduce_partial");
    if (rc) return rc;
    hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(64), 0, h->stream, h->d_part, h->red_blocks,
                       h->d_out, (double)h->g.nx * h->g.ny * h->g.nz);
    if ((rc = last_launch(h, "k_reduce_final"))) return rc;
    HIPCHK(h, hipMemcpyAsync(out, h->d_out, 4 * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return TFF_OK;
  }
};

// After tff_init_fields / tff_init_spectrum wrote the (clean) state and B0
// in place, as an upload of every field would.
template <class S>
int after_init(tff_handle* h, const char* what) {
  int rc = last_launch(h, what);
  if (rc) return rc;
  HIPCHK(h, hipStreamSynchronize(h->stream));
  h->ghosts_settled = false;
  h->t0_dirty = true;
  return TFF_OK;
}

template <class PR>
struct DoInitFields {
  static int run(tff_handle* h, int kind, const double* params, int nparams) {
    using S = typename PR::S;
    double* dp = nullptr;
    HIPCHK(h, hipStreamSynchronize(h->stream));
    if (hipMalloc(&dp, 16 * sizeof(double)) != hipSuccess) return fail(h, TFF_ENOMEM, "tff_init_fields");
    double p[16] = {0};
    for (int q = 0; q < nparams; ++q) p[q] = params[q];
    int rc = TFF_OK;
    if (hipMemcpyAsync(dp, p, sizeof p, hipMemcpyHostToDevice, h->stream) != hipSuccess)
      rc = fail(h, TFF_EHIP, "tff_init_fields: parameters");
    const int64_t n = (int64_t)(h->g.nx + 2 * NG) * (h->g.ny + 2 * NG) * (h->g.nz + 2 * NG);
    const unsigned nb = (unsigned)((n + 255) / 256);
    if (!rc) {
      auto e = mv3<S>(h->es), b = mv3<S>(h->bs), b0 = mv3<S>(h->b0);
      const int64_t lo = h->comp ? 3 * h->g.cs : 0;
      if (kind == TFF_IC_LOOP)
        hipLaunchKernelGGL(k_init_loop<S>, dim3(nb), dim3(256), 0, h->stream, h->g, e, b, b0, dp, n, lo);
      else if (kind == TFF_IC_PSR)
        hipLaunchKernelGGL(k_init_psr<S>, dim3(nb), dim3(256), 0, h->stream, h->g, e, b, b0, dp, n, lo);
      else
        hipLaunchKernelGGL(k_init_uniform<S>, dim3(nb), dim3(256), 0, h->stream, h->g, e, b, b0, dp, n, lo);
      rc = after_init<S>(h, "k_init_fields");
    }
    hipFree(dp);
    return rc;
  }
};

template <class PR>
struct DoInitSpectrum {
  static int run(tff_handle* h, const double* modes, int nmodes, const double* background) {
    using S = typename PR::S;
    double* dm = nullptr;
    const size_t mb = ((size_t)nmodes * SPEC_MODE + 6) * sizeof(double);
    HIPCHK(h, hipStreamSynchronize(h->stream));
    if (hipMalloc(&dm, mb) != hipSuccess) return fail(h, TFF_ENOMEM, "tff_init_spectrum");
    int rc = TFF_OK;
    if (hipMemcpyAsync(dm, background, 6 * sizeof(double), hipMemcpyHostToDevice, h->stream) != hipSuccess ||
        (nmodes > 0 && hipMemcpyAsync(dm + 6, modes, (size_t)nmodes * SPEC_MODE * sizeof(double),
                                      hipMemcpyHostToDevice, h->stream) != hipSuccess))
      rc = fail(h, TFF_EHIP, "tff_init_spectrum: modes");
    const int64_t n = (int64_t)(h->g.nx + 2 * NG) * (h->g.ny + 2 * NG) * (h->g.nz + 2 * NG);
    if (!rc) {
      hipLaunchKernelGGL(k_init_spectrum<S>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, h->stream, h->g,
                         mv3<S>(h->es), mv3<S>(h->bs), mv3<S>(h->b0), (const double*)(dm + 6), nmodes,
                         (const double*)dm, n, h->comp ? 3 * h->g.cs : (int64_t)0);
      rc = after_init<S>(h, "k_init_spectrum");
    }
    hipFree(dm);
    return rc;
  }
};

template <class PR>
struct TileLds {
  static int run(tff_handle* h, int t, size_t* out) {
    *out = tile_lds<PR>(t);
    return TFF_OK;
  }
};

int ensure_clean(tff_handle* h) { return dispatch<DoEnsureClean>(h); }

Vec* field_vec(tff_handle* h, int field, int& comp) {
  comp = field % 3;
  switch (field / 3) {
    case 0: return &h->es;
    case 1: return &h->bs;
    case 2: return &h->b0;
    case 3: return &h->en;
    case 4: return &h->bn;
    default: return nullptr;
  }
}

}  // namespace

// ---- staging without allocation ------------------------------------------
// tff_output_fields_noalloc stages through `scratch`, a file-scope device
// array that every GPU the library runs on holds (one copy each) until the
// process ends.  In host code `scratch` names a host-side stand-in, never
// the device memory: each device's own address is looked up once, with that
// device current (tff_create), and only that address is used afterwards.
static __device__ __align__(256) unsigned char scratch[TFF_NOALLOC_SCRATCH_BYTES];
namespace {
std::mutex scratch_mu;
std::map<int, void*> scratch_addr;  // device -> its copy of scratch

int scratch_setup(int device) {
  std::lock_guard<std::mutex> lock(scratch_mu);
  if (scratch_addr.count(device)) return TFF_OK;
  void* p = nullptr;
  if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return TFF_EHIP;
  scratch_addr[device] = p;
  return TFF_OK;
}

void* scratch_of(int device) {
  std::lock_guard<std::mutex> lock(scratch_mu);
  auto it = scratch_addr.find(device);
  return it == scratch_addr.end() ? nullptr : it->second;
}
}  // namespace

// ============================================================================
extern "C" {

int tff_create(const tff_config* cfg, tff_handle** out) {
  if (!cfg || !out) return TFF_EINVAL;
  *out = nullptr;
  if (cfg->nx < NG || cfg->ny < NG || cfg->nz < NG) return TFF_EINVAL;  // initialize.F90:256-264
  if (cfg->precision < TFF_FP64 || cfg->precision > TFF_FP32_COMP) return TFF_EINVAL;
  for (int a = 0; a < 3; ++a)
    if (cfg->boundary[a] < 0 || cfg->boundary[a] > 2) return TFF_EINVAL;
  if (!(cfg->cc > 0.0) || cfg->absorb_buff < 0) return TFF_EINVAL;
  tff_handle* h = new tff_handle();
  h->cfg = *cfg;
  // initialize.F90:143-148: one absorbing axis makes all three absorbing
  if (cfg->boundary[0] == TFF_ABSORB || cfg->boundary[1] == TFF_ABSORB || cfg->boundary[2] == TFF_ABSORB)
    h->cfg.boundary[0] = h->cfg.boundary[1] = h->cfg.boundary[2] = TFF_ABSORB;
  h->esz = cfg->precision == TFF_FP64 ? 8 : 4;
  h->comp = cfg->precision == TFF_FP32_COMP;
  if (h->comp) h->fused = 0;
  const int al = (int)(128 / h->esz);  // elements per 128-byte line
  h->xo = al;
  h->g.nx = cfg->nx;
  h->g.ny = cfg->ny;
  h->g.nz = cfg->nz;
  h->g.px = ((h->xo + cfg->nx + NG + al - 1) / al) * al;
  h->g.plane = (int64_t)h->g.px * (cfg->ny + 2 * NG);
  h->g.org = h->xo + (int64_t)NG * h->g.px + (int64_t)NG * h->g.plane;
  h->nalloc = h->g.plane * (cfg->nz + 2 * NG) + al;
  h->nalloc = (h->nalloc + 63) / 64 * 64;  // keep every component 256-byte aligned
  h->g.cs = h->nalloc;
  if (hipSetDevice(cfg->device) != hipSuccess || scratch_setup(cfg->device) != TFF_OK) {
    delete h;
    return TFF_EHIP;
  }
  {
    int lds = 0;
    if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, cfg->device) == hipSuccess &&
        lds > 0)
      h->max_lds = lds;
  }
  if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
    delete h;
    return TFF_EHIP;
  }
  int rc = TFF_OK;
  Vec* vs[] = {&h->es, &h->bs, &h->w1e, &h->w1b, &h->w2e, &h->w2b, &h->ce, &h->b0, &h->ge};
  for (Vec* v : vs)
    if ((rc = alloc_vec(h, *v))) break;
  h->red_blocks = 1024;
  if (!rc && (hipMalloc(&h->d_step, sizeof(int)) != hipSuccess ||
              hipMalloc(&h->d_flag, sizeof(int)) != hipSuccess ||
              hipMalloc(&h->d_bc, 16 * sizeof(double)) != hipSuccess ||
              hipMalloc(&h->d_part, 4 * sizeof(double) * h->red_blocks) != hipSuccess ||
              hipMalloc(&h->d_out, 4 * sizeof(double)) != hipSuccess))
    rc = TFF_ENOMEM;
  if (!rc && hipMemsetAsync(h->d_bc, 0, 16 * sizeof(double), h->stream) != hipSuccess) rc = TFF_EHIP;
  if (!rc && cfg->precision == TFF_FP32_BULK) {
    // the most tiles any tile shape and z-chunk can make (32 x 4, one plane)
    h->fix_cap = (int64_t)((cfg->nx + 31) / 32) * ((cfg->ny + 3) / 4) * cfg->nz;
    if (hipMalloc(&h->d_fixmask, (size_t)h->fix_cap * sizeof(unsigned long long)) != hipSuccess) rc = TFF_ENOMEM;
  }
  if (!rc && hipStreamSynchronize(h->stream) != hipSuccess) rc = TFF_EHIP;
  for (int s = 0; !rc && s < 16; ++s)
    if (hipEventCreate(&h->ev[s]) != hipSuccess) rc = TFF_EHIP;
  if (rc) {
    tff_destroy(h);
    return rc;
  }
  *out = h;
  return TFF_OK;
}

void tff_destroy(tff_handle* h) {
  if (!h) return;
  if (h->stream) hipStreamSynchronize(h->stream);
  invalidate_graphs(h);
  drop_retired(h);
  for (void* p : h->allocs) hipFree(p);
  if (h->out_buf) hipFree(h->out_buf);
  if (h->psr_stage) hipFree(h->psr_stage);
  hipFree(h->d_step);
  hipFree(h->d_flag);
  if (h->d_fixmask) hipFree(h->d_fixmask);
  hipFree(h->d_bc);
  hipFree(h->d_part);
  hipFree(h->d_out);
  for (auto e : h->ev)
    if (e) hipEventDestroy(e);
  for (auto e : h->tpool) hipEventDestroy(e);
  for (auto e : h->ppool) hipEventDestroy(e);
  if (h->stream) hipStreamDestroy(h->stream);
  delete h;
}

const char* tff_last_error(const tff_handle* h) { return h ? h->err.c_str() : "null handle"; }

// Transfers of a compensated handle: an FP64 upload splits each value into
// hi = fl32(x) and lo = fl32(x - hi), an FP32 upload sets lo = 0; an FP64
// download returns hi + lo (exact in FP64), an FP32 download hi.
static int xfer_pairs(tff_handle* h, void* buf, int comp, void* host, bool f64, bool up, int sx, int sy, int sz) {
  const size_t n = (size_t)sx * sy * sz;
  std::vector<float> hi(n), lo(n);
  char* dhi = (char*)buf + ((size_t)comp * h->g.cs + (size_t)(h->xo - NG)) * 4;
  char* dlo = dhi + (size_t)3 * h->g.cs * 4;
  const size_t pitch = (size_t)h->g.px * 4, row = (size_t)sx * 4;
  HIPCHK(h, hipStreamSynchronize(h->stream));
  if (up) {
    for (size_t t = 0; t < n; ++t) {
      const double x = f64 ? ((const double*)host)[t] : (double)((const float*)host)[t];
      hi[t] = (float)x;
      lo[t] = (float)(x - (double)hi[t]);
    }
    HIPCHK(h, hipMemcpy2DAsync(dhi, pitch, hi.data(), row, row, (size_t)sy * sz, hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipMemcpy2DAsync(dlo, pitch, lo.data(), row, row, (size_t)sy * sz, hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return TFF_OK;
  }
  HIPCHK(h, hipMemcpy2DAsync(hi.data(), row, dhi, pitch, row, (size_t)sy * sz, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipMemcpy2DAsync(lo.data(), row, dlo, pitch, row, (size_t)sy * sz, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  for (size_t t = 0; t < n; ++t) {
    if (f64)
      ((double*)host)[t] = (double)hi[t] + (double)lo[t];
    else
      ((float*)host)[t] = hi[t];
  }
  return TFF_OK;
}

static int xfer(tff_handle* h, int field, void* host, int host_dtype, bool up) {
  if (!h || !host) return TFF_EINVAL;
  if (field < TFF_EX || field > TFF_BNZ) return fail(h, TFF_EINVAL, "unknown field id");
  if (host_dtype != TFF_HOST_F64 && host_dtype != TFF_HOST_F32)
    return fail(h, TFF_EINVAL, "unknown host dtype");
  int rc0 = ensure_clean(h);
  if (rc0) return rc0;
  if (up) h->ghosts_settled = false;  // new ghost values until a full step carries them
  int comp;
  Vec* v = field_vec(h, field, comp);
  if (field >= TFF_ENX && !v->p[0]) {
    if (!up) return fail(h, TFF_ESTATE, "En/Bn are not allocated before tff_init_rk_step");
    int rc = alloc_vec(h, *v);
    if (rc) return rc;
  }
  const int sx = h->g.nx + 2 * NG, sy = h->g.ny + 2 * NG, sz = h->g.nz + 2 * NG;
  const size_t n = (size_t)sx * sy * sz;
  const size_t hsz = host_dtype == TFF_HOST_F64 ? 8 : 4;
  if (h->comp) return xfer_pairs(h, v->p[0], comp, host, hsz == 8, up, sx, sy, sz);
  std::vector<unsigned char> tmp;
  void* hp = host;
  if (hsz != h->esz) {
    tmp.resize(n * h->esz);
    hp = tmp.data();
    if (up) {
      if (h->esz == 4)
        for (size_t t = 0; t < n; ++t) ((float*)hp)[t] = (float)((const double*)host)[t];
      else
        for (size_t t = 0; t < n; ++t) ((double*)hp)[t] = (double)((const float*)host)[t];
    }
  }
  char* dbase = (char*)v->p[comp] + (size_t)(h->xo - NG) * h->esz;
  HIPCHK(h, hipStreamSynchronize(h->stream));
  if (up)
    HIPCHK(h, hipMemcpy2DAsync(dbase, (size_t)h->g.px * h->esz, hp, (size_t)sx * h->esz, (size_t)sx * h->esz,
                               (size_t)sy * sz, hipMemcpyHostToDevice, h->stream));
  else
    HIPCHK(h, hipMemcpy2DAsync(hp, (size_t)sx * h->esz, dbase, (size_t)h->g.px * h->esz, (size_t)sx * h->esz,
                               (size_t)sy * sz, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));  // the host buffer is in use until here
  if (!up && hsz != h->esz) {
    if (h->esz == 4)
      for (size_t t = 0; t < n; ++t) ((double*)host)[t] = (double)((const float*)hp)[t];
    else
      for (size_t t = 0; t < n; ++t) ((float*)host)[t] = (float)((const double*)hp)[t];
  }
  if (up && field >= TFF_B0X && field <= TFF_B0Z) h->t0_dirty = true;
  return TFF_OK;
}

int tff_upload(tff_handle* h, int field, const void* host, int host_dtype) {
  return xfer(h, field, const_cast<void*>(host), host_dtype, true);
}

int tff_download(tff_handle* h, int field, void* host, int host_dtype) {
  return xfer(h, field, host, host_dtype, false);
}

int tff_set_bc(tff_handle* h, int kind, const double* params, int nparams) {
  if (!h) return TFF_EINVAL;
  if (kind < TFF_BC_NONE || kind > TFF_BC_UNIFORM) return fail(h, TFF_EINVAL, "unknown BC kind");
  static const int need[4] = {0, 9, 6, 6};
  if (nparams < need[kind] || nparams > 16 || (nparams > 0 && !params))
    return fail(h, TFF_EINVAL, "wrong number of BC parameters");
  int rc0 = ensure_clean(h);
  if (rc0) return rc0;
  std::memset(h->bc_host, 0, sizeof(h->bc_host));
  for (int t = 0; t < nparams; ++t) h->bc_host[t] = params[t];
  HIPCHK(h, hipStreamSynchronize(h->stream));
  // captured graphs hold the hook's staging buffer and box: never replay
  // them with the new parameters (on every path, whatever the kind)
  invalidate_graphs(h);
  drop_retired(h);
  HIPCHK(h, hipMemcpyAsync(h->d_bc, h->bc_host, 16 * sizeof(double), hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  if (h->psr_stage) {
    hipFree(h->psr_stage);
    h->psr_stage = nullptr;
  }
  h->psr_n = int3{0, 0, 0};
  if (kind == TFF_BC_PSR) {
    // bounding box of the supersphere r < 2R (user_psr.F90:91,153), clipped
    // to the interior the BC loops over
    const double R2 = 2.0 * params[0];
    const double c[3] = {params[3], params[4], params[5]};
    const int n[3] = {h->g.nx, h->g.ny, h->g.nz};
    int lo[3], hi[3];
    for (int a = 0; a < 3; ++a) {
      lo[a] = std::max(0, (int)std::floor(c[a] - R2) - 1);
      hi[a] = std::min(n[a] - 1, (int)std::ceil(c[a] + R2) + 1);
    }
    if (hi[0] >= lo[0] && hi[1] >= lo[1] && hi[2] >= lo[2]) {
      h->psr_lo = int3{lo[0], lo[1], lo[2]};
      h->psr_n = int3{hi[0] - lo[0] + 1, hi[1] - lo[1] + 1, hi[2] - lo[2] + 1};
      const size_t nb = (size_t)h->psr_n.x * h->psr_n.y * h->psr_n.z * 6 * h->esz;
      if (hipMalloc(&h->psr_stage, nb) != hipSuccess) return fail(h, TFF_ENOMEM, "psr staging");
    }
  }
  h->bc_kind = kind;
  return TFF_OK;
}

int tff_init_rk_step(tff_handle* h) {
  if (!h) return TFF_EINVAL;
  {
    int rc0 = ensure_clean(h);
    if (rc0) return rc0;
  }
  for (Vec* v : {&h->en, &h->bn})
    if (!v->p[0]) {
      int rc = alloc_vec(h, *v);
      if (rc) return rc;
    }
  const int64_t n = h->nalloc;
  const int tb = 256;
  if (h->esz == 8) {
    hipLaunchKernelGGL(k_copy3<double>, dim3((unsigned)((n + tb - 1) / tb)), dim3(tb), 0, h->stream, n,
                       cv3<double>(h->es), mv3<double>(h->en));
    hipLaunchKernelGGL(k_copy3<double>, dim3((unsigned)((n + tb - 1) / tb)), dim3(tb), 0, h->stream, n,
                       cv3<double>(h->bs), mv3<double>(h->bn));
  } else {
    // (the compensated pairs: the lo words follow as three more components)
    for (int64_t lo = 0; lo <= (h->comp ? 3 * h->g.cs : 0); lo += 3 * h->g.cs) {
      auto sh = [lo](V3<float> v) { return V3<float>{v.x + lo, v.y + lo, v.z + lo}; };
      auto shc = [lo](V3<const float> v) { return V3<const float>{v.x + lo, v.y + lo, v.z + lo}; };
      hipLaunchKernelGGL(k_copy3<float>, dim3((unsigned)((n + tb - 1) / tb)), dim3(tb), 0, h->stream, n,
                         shc(cv3<float>(h->es)), sh(mv3<float>(h->en)));
      hipLaunchKernelGGL(k_copy3<float>, dim3((unsigned)((n + tb - 1) / tb)), dim3(tb), 0, h->stream, n,
                         shc(cv3<float>(h->bs)), sh(mv3<float>(h->bn)));
    }
  }
  return last_launch(h, "k_copy3");
}

int tff_rk3_step(tff_handle* h, double c1, double c2, double c3) {
  if (!h) return TFF_EINVAL;
  {
    int rc0 = ensure_clean(h);
    if (rc0) return rc0;
  }
  if (!h->en.p[0]) return fail(h, TFF_ESTATE, "tff_rk3_step before tff_init_rk_step");
  return dispatch<DoRk3>(h, c1, c2, c3);
}

int tff_clean_epar(tff_handle* h) {
  if (!h) return TFF_EINVAL;
  {
    int rc0 = ensure_clean(h);
    if (rc0) return rc0;
  }
  int rc = dispatch<DoClean>(h);
  invalidate_graphs(h);
  return rc;
}

int tff_exchange_fields(tff_handle* h, int exchange_b0) {
  if (!h) return TFF_EINVAL;
  {
    int rc0 = ensure_clean(h);
    if (rc0) return rc0;
  }
  return dispatch<DoFill>(h, exchange_b0);
}

int tff_user_bc(tff_handle* h, int step) {
  if (!h) return TFF_EINVAL;
  {
    int rc0 = ensure_clean(h);
    if (rc0) return rc0;
  }
  return dispatch<DoBC>(h, step);
}

int tff_step(tff_handle* h, int timestep) {
  if (!h) return TFF_EINVAL;
  int done = 0;
  return dispatch<DoStep>(h, timestep, 0, 1, &done);
}

int tff_run(tff_handle* h, int first_step, int nsteps) {
  if (!h || nsteps < 0) return TFF_EINVAL;
  if (nsteps == 0) return TFF_OK;
  hipLaunchKernelGGL(k_set_int, dim3(1), dim3(64), 0, h->stream, h->d_step, first_step);
  HIPCHK(h, hipStreamSynchronize(h->stream));
  for (int n = 0; n < nsteps;) {
    int done = 0;
    const int count = nsteps - n >= RUN_GRAPH_STEPS ? RUN_GRAPH_STEPS : 1;
    int rc = dispatch<DoStep>(h, first_step + n, 1, count, &done);
    if (rc) return rc;
    n += done;
  }
  return TFF_OK;
}

int tff_output_dims(tff_handle* h, int istep, int dims[3]) {
  if (!h || !dims) return TFF_EINVAL;
  if (istep < 2)
    return fail(h, TFF_EINVAL, "output_istep must be >= 2 (the istep = 1 debug layout is not supported)");
  return output_dims(h, istep, dims);
}

int tff_output_fields(tff_handle* h, int istep, void* const out[6], int host_dtype) {
  if (!h || !out) return TFF_EINVAL;
  if (istep < 2)
    return fail(h, TFF_EINVAL, "output_istep must be >= 2 (the istep = 1 debug layout is not supported)");
  if (host_dtype != TFF_HOST_F64 && host_dtype != TFF_HOST_F32) return fail(h, TFF_EINVAL, "unknown host dtype");
  for (int f = 0; f < 6; ++f)
    if (!out[f]) return fail(h, TFF_EINVAL, "null output buffer");
  int rc = ensure_clean(h);
  if (rc) return rc;
  const size_t need = output_bytes(h, istep, host_dtype);
  if (need > h->out_cap) {
    HIPCHK(h, hipStreamSynchronize(h->stream));
    if (h->out_buf) hipFree(h->out_buf);
    h->out_buf = nullptr;
    h->out_cap = 0;
    if (hipMalloc(&h->out_buf, need) != hipSuccess) return fail(h, TFF_ENOMEM, "hipMalloc of the output buffer failed");
    h->out_cap = need;
  }
  return dispatch<DoOutput>(h, istep, out, host_dtype, h->out_buf);
}
int tff_output_samples(tff_handle* h, int istep, void* const out[7], int host_dtype) {
  if (!h || !out) return TFF_EINVAL;
  if (istep < 2)
    return fail(h, TFF_EINVAL, "output_istep must be >= 2 (the istep = 1 debug layout is not supported)");
  if (host_dtype != TFF_HOST_F64 && host_dtype != TFF_HOST_F32) return fail(h, TFF_EINVAL, "unknown host dtype");
  for (int f = 0; f < 7; ++f)
    if (!out[f]) return fail(h, TFF_EINVAL, "null output buffer");
  int rc = ensure_clean(h);
  if (rc) return rc;
  const size_t need = output_bytes(h, istep, host_dtype) / 6 * 7;
  if (need > h->out_cap) {
    HIPCHK(h, hipStreamSynchronize(h->stream));
    if (h->out_buf) hipFree(h->out_buf);
    h->out_buf = nullptr;
    h->out_cap = 0;
    if (hipMalloc(&h->out_buf, need) != hipSuccess) return fail(h, TFF_ENOMEM, "hipMalloc of the output buffer failed");
    h->out_cap = need;
  }
  return dispatch<DoSamples>(h, istep, out, host_dtype, h->out_buf);
}
int tff_output_fields_noalloc(tff_handle* h, int istep, void* const out[6], int host_dtype) {
  if (!h || !out) return TFF_EINVAL;
  if (istep < 2)
    return fail(h, TFF_EINVAL, "output_istep must be >= 2 (the istep = 1 debug layout is not supported)");
  if (host_dtype != TFF_HOST_F64 && host_dtype != TFF_HOST_F32) return fail(h, TFF_EINVAL, "unknown host dtype");
  for (int f = 0; f < 6; ++f)
    if (!out[f]) return fail(h, TFF_EINVAL, "null output buffer");
  const size_t need = output_bytes(h, istep, host_dtype);
  if (need > (size_t)TFF_NOALLOC_SCRATCH_BYTES)
    return fail(h, TFF_EINVAL,
                "tff_output_fields_noalloc: the six datasets take " + std::to_string(need) +
                    " bytes, more than its " + std::to_string(TFF_NOALLOC_SCRATCH_BYTES) +
                    "-byte scratch (use tff_output_fields, or a larger istep)");
  void* p = scratch_of(h->cfg.device);
  if (!p) return fail(h, TFF_ESTATE, "no scratch address for this device");
  // cleaning the state only launches kernels (invalidated graphs are freed
  // later, where the library captures anew)
  int rc = ensure_clean(h);
  if (rc) return rc;
  return dispatch<DoOutput>(h, istep, out, host_dtype, p);
}

int tff_reduce(tff_handle* h, double out[4]) {
  if (!h || !out) return TFF_EINVAL;
  {
    int rc0 = ensure_clean(h);
    if (rc0) return rc0;
  }
  return dispatch<DoReduce>(h, out);
}

int tff_sync(tff_handle* h) {
  if (!h) return TFF_EINVAL;
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return TFF_OK;
}

int tff_event_record(tff_handle* h, int slot) {
  if (!h || slot < 0 || slot >= 16) return TFF_EINVAL;
  HIPCHK(h, hipEventRecord(h->ev[slot], h->stream));
  return TFF_OK;
}

int tff_event_elapsed(tff_handle* h, int a, int b, float* ms) {
  if (!h || a < 0 || a >= 16 || b < 0 || b >= 16 || !ms) return TFF_EINVAL;
  HIPCHK(h, hipEventSynchronize(h->ev[b]));
  HIPCHK(h, hipEventElapsedTime(ms, h->ev[a], h->ev[b]));
  return TFF_OK;
}

int tff_set_kernel_timing(tff_handle* h, int on) {
  if (!h) return TFF_EINVAL;
  HIPCHK(h, hipStreamSynchronize(h->stream));
  h->timing = on != 0;
  h->tused = 0;
  h->tsum_ms = 0.0;
  h->tcount = 0;
  return TFF_OK;
}

int tff_set_path(tff_handle* h, int fused) {
  if (!h) return TFF_EINVAL;
  if (h->comp && fused) return fail(h, TFF_EINVAL, "TFF_FP32_COMP runs the per-procedure kernels only");
  int rc = ensure_clean(h);
  if (rc) return rc;
  h->fused = fused != 0;
  invalidate_graphs(h);
  return TFF_OK;
}

int tff_set_tuning(tff_handle* h, int tile, int kchunk) {
  if (!h) return TFF_EINVAL;
  if (tile < -1 || tile >= NTILES || (tile >= 0 && (kchunk < 4 || kchunk > 64 || kchunk % 4 != 0)))
    return fail(h, TFF_EINVAL, "bad tuning (tile index out of range, or z-chunk not a multiple of 4 in 4..64)");
  if (tile >= 0) {
    size_t lds = 0;
    dispatch<TileLds>(h, tile, &lds);
    if (lds > (size_t)h->max_lds) return fail(h, TFF_EINVAL, "tile shape exceeds the device's LDS per block");
  }
  HIPCHK(h, hipStreamSynchronize(h->stream));
  h->tile = tile;
  h->kchunk = tile >= 0 ? kchunk : 0;
  invalidate_graphs(h);
  return TFF_OK;
}

int tff_get_tuning(tff_handle* h, int* tile, int* tx, int* ty, int* kchunk) {
  if (!h || !tile || !tx || !ty || !kchunk) return TFF_EINVAL;
  *tile = h->tile;
  *tx = h->tile >= 0 ? TILE_TX[h->tile] : 0;
  *ty = h->tile >= 0 ? TILE_TY[h->tile] : 0;
  *kchunk = h->kchunk;
  return TFF_OK;
}

int tff_init_fields(tff_handle* h, int kind, const double* params, int nparams) {
  if (!h || (nparams > 0 && !params)) return TFF_EINVAL;
  static const int need[4] = {0, 5, 6, 4};
  if (kind < TFF_IC_LOOP || kind > TFF_IC_UNIFORM || nparams < need[kind] || nparams > 16)
    return fail(h, TFF_EINVAL, "tff_init_fields: unknown kind or wrong number of parameters");
  int rc = ensure_clean(h);  // the buffers then hold the reference's roles
  if (rc) return rc;
  return dispatch<DoInitFields>(h, kind, params, nparams);
}
int tff_init_spectrum(tff_handle* h, const double* modes, int nmodes, const double background[6]) {
  if (!h || nmodes < 0 || (nmodes > 0 && !modes) || !background) return TFF_EINVAL;
  const double kcap = 2147483647.0 / (std::max(h->g.nx, std::max(h->g.ny, h->g.nz)) + 2 * NG);
  for (int m = 0; m < nmodes; ++m)
    for (int a = 0; a < 3; ++a) {
      const double k = modes[(size_t)m * SPEC_MODE + a];
      if (!(std::fabs(k) < kcap) || k != std::floor(k))
        return fail(h, TFF_EINVAL, "tff_init_spectrum: mode " + std::to_string(m) + ": wave number " +
                                       std::to_string(k) + " is not an integer below " + std::to_string(kcap));
    }
  int rc = ensure_clean(h);
  if (rc) return rc;
  return dispatch<DoInitSpectrum>(h, modes, nmodes, background);
}
int tff_get_size(tff_handle* h, int n[3], int* precision) {
  if (!h || !n) return TFF_EINVAL;
  n[0] = h->g.nx;
  n[1] = h->g.ny;
  n[2] = h->g.nz;
  if (precision) *precision = h->cfg.precision;
  return TFF_OK;
}
int tff_get_layout(tff_handle* h, long long out[5]) {
  if (!h || !out) return TFF_EINVAL;
  out[0] = h->g.px;
  out[1] = h->g.plane;
  out[2] = h->g.org;
  out[3] = h->nalloc;
  out[4] = (long long)h->esz;
  return TFF_OK;
}

int tff_set_phase_timing(tff_handle* h, int on) {
  if (!h) return TFF_EINVAL;
  HIPCHK(h, hipStreamSynchronize(h->stream));
  h->phases = on != 0;
  h->pused = 0;
  for (double& v : h->psum_ms) v = 0.0;
  return TFF_OK;
}
int tff_phase_times(tff_handle* h, double ms[TFF_PHASES]) {
  if (!h || !ms) return TFF_EINVAL;
  HIPCHK(h, hipStreamSynchronize(h->stream));
  for (size_t t = 0; t + 1 < h->pused; t += 2) {
    float m = 0.f;
    HIPCHK(h, hipEventElapsedTime(&m, h->ppool[t], h->ppool[t + 1]));
    h->psum_ms[h->pphase[t / 2]] += m;
  }
  h->pused = 0;
  for (int p = 0; p < TFF_PHASES; ++p) {
    ms[p] = h->psum_ms[p];
    h->psum_ms[p] = 0.0;
  }
  return TFF_OK;
}
int tff_stage_kernel_time(tff_handle* h, double* total_ms, long* launches) {
  if (!h || !total_ms || !launches) return TFF_EINVAL;
  HIPCHK(h, hipStreamSynchronize(h->stream));
  for (size_t t = 0; t + 1 < h->tused; t += 2) {
    float ms = 0.f;
    HIPCHK(h, hipEventElapsedTime(&ms, h->tpool[t], h->tpool[t + 1]));
    h->tsum_ms += ms;
    h->tcount += 1;
  }
  h->tused = 0;
  *total_ms = h->tsum_ms;
  *launches = h->tcount;
  return TFF_OK;
}

}  // extern "C"

// ============================================================================
// Batched path: independent meshblocks advanced together, one persistent
// queue-driven launch of the fused stage kernel per RK stage.

// block size of the batched ghost passes (their per-instance lists are
// padded to it)
constexpr int BATCH_TB = 256;

struct tff_batch {
  int precision = TFF_FP32_BULK;
  int device = 0;
  std::vector<tff_handle*> inst;
  hipStream_t stream = nullptr;
  std::string err;
  int tile = -1, kchunk = 0, grid = 0;
  int64_t total_tiles = 0, total_shell = 0;
  int nghost = 0;
  bool first = true;
  int64_t* d_tile_base = nullptr;
  int64_t* d_shell_base = nullptr;
  int* d_shell_inst = nullptr;
  // ghost-clean list (stored-ghost instances) and edge-ring list (WRAP_MIXED
  // instances); the shell list above feeds the build pass of both kinds
  int64_t* d_clean_base = nullptr;
  int* d_clean_inst = nullptr;
  int nclean = 0;
  int64_t total_clean = 0;
  int64_t* d_edge_base = nullptr;
  int* d_edge_inst = nullptr;
  int nedge = 0;
  int64_t total_edge = 0;
  int64_t* d_floor_base = nullptr;  // driven-loop instances: interior floor planes
  int* d_floor_inst = nullptr;
  int nfloor = 0;
  int64_t total_floor = 0;
  void* d_info = nullptr;
  void* d_stage = nullptr;  // BatchStage[4][n]: stage 1 from a clean state, stages 1, 2, 3
  int* d_step = nullptr;
  void* d_fixmask = nullptr;  // FP32-bulk: one word per work-list tile
  unsigned* d_queue = nullptr;  // k_batch_stage's segment counters (cleared by each launch)
  BatchQueue qseg{};            // the work list cut into 8 segments of about equal cost
  hipGraphExec_t graph = nullptr;
};

namespace {

int bfail(tff_batch* b, int code, const std::string& msg) {
  if (b) b->err = msg;
  return code;
}

#define BHIP(b, call)                                                                     \
  do {                                                                                    \
    hipError_t e_ = (call);                                                               \
    if (e_ != hipSuccess)                                                                 \
      return bfail((b), TFF_EHIP, std::string(#call) + ": " + hipGetErrorString(e_));     \
  } while (0)

template <class PR>
int batch_tile_lds_ok(tff_batch* b, int t) {
  return tile_lds<PR>(t) <= (size_t)b->inst[0]->max_lds;
}

template <class PR, class TL>
int batch_launch_stage_t(tff_batch* b, int which) {
  using S = typename PR::S;
  using T = typename PR::T;
  using P = typename PR::P;
  if constexpr (!tile_fits<PR, TL>()) return bfail(b, TFF_EINVAL, "tile shape exceeds the LDS of a CU");
  else {
  const size_t lds = stage_lds<PR, TL, PR::null_fp64>();
  static bool attr = false;
  if (!attr) {
    hipFuncSetAttribute((const void*)k_batch_stage<PR, TL>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    attr = true;
  }
  const int n = (int)b->inst.size();
  const BatchStage<S, P>* st = (const BatchStage<S, P>*)b->d_stage + (size_t)which * n;
  const unsigned grid = (unsigned)std::min<int64_t>(b->grid, b->total_tiles);
  BatchQueue q = b->qseg;
  if (grid % 8 != 0) {  // one queue: segment 0 holds every item
    for (int x = 1; x < 9; ++x) q.seg[x] = b->total_tiles;
    q.seg[0] = 0;
  }
  hipLaunchKernelGGL((k_batch_stage<PR, TL>), dim3(grid), dim3(TL::NT), lds, b->stream, st, b->d_tile_base, n, q,
                     b->d_queue);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return bfail(b, TFF_EHIP, std::string("k_batch_stage: ") + hipGetErrorString(e));
  if constexpr (PR::null_fp64) {
    const size_t flds = stage_lds<PR, TL, false>();
    static bool fattr = false;
    if (!fattr) {
      hipFuncSetAttribute((const void*)k_batch_stage_fix<PR, TL>, hipFuncAttributeMaxDynamicSharedMemorySize,
                          (int)flds);
      fattr = true;
    }
    const unsigned fgrid = (unsigned)std::min<int64_t>(FIX_GRID, b->total_tiles);
    hipLaunchKernelGGL((k_batch_stage_fix<PR, TL>), dim3(fgrid), dim3(TL::NT), flds, b->stream, st,
                       b->d_tile_base, n, b->total_tiles, (const unsigned long long*)b->d_fixmask);
    e = hipGetLastError();
    if (e != hipSuccess) return bfail(b, TFF_EHIP, std::string("k_batch_stage_fix: ") + hipGetErrorString(e));
  }
  return TFF_OK;
  }
}

template <class PR>
int batch_launch_stage(tff_batch* b, int which) {
  return with_tile<int>(b->tile, [&](auto tl) { return batch_launch_stage_t<PR, decltype(tl)>(b, which); });
}

template <class PR, class TL>
int batch_grid_t(tff_batch* b) {
  using S = typename PR::S;
  using T = typename PR::T;
  if constexpr (!tile_fits<PR, TL>()) return bfail(b, TFF_EINVAL, "tile shape exceeds the LDS of a CU");
  else {
  const size_t lds = stage_lds<PR, TL, PR::null_fp64>();
  hipFuncSetAttribute((const void*)k_batch_stage<PR, TL>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  int per_cu = 0;
  BHIP(b, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)k_batch_stage<PR, TL>, TL::NT, lds));
  hipDeviceProp_t prop;
  BHIP(b, hipGetDeviceProperties(&prop, b->device));
  b->grid = std::max(1, per_cu) * prop.multiProcessorCount;
  return TFF_OK;
  }
}

template <class PR>
int batch_grid(tff_batch* b) {
  return with_tile<int>(b->tile, [&](auto tl) { return batch_grid_t<PR, decltype(tl)>(b); });
}

// Per-instance stage descriptors and work lists for the current tile shape.
template <class PR, class TL>
int batch_build_t(tff_batch* b) {
  using S = typename PR::S;
  using P = typename PR::P;
  const int n = (int)b->inst.size();
  std::vector<BatchStage<S, P>> st(4 * (size_t)n);
  std::vector<int64_t> tile_base(n), shell_base, floor_base, clean_base, edge_base;
  std::vector<int> shell_inst, floor_inst, clean_inst, edge_inst;
  std::vector<BatchInfo> info(n);
  std::vector<double> cost;
  int64_t tt = 0, ss = 0, ff = 0, cl = 0, ee = 0;
  for (int q = 0; q < n; ++q) {
    tff_handle* h = b->inst[q];
    h->kchunk = b->kchunk;
    const bool wrap = wrap_mode(h), mixed = mixed_mode(h);
    auto a = stage_common<PR>(h);
    a.nbx = (h->g.nx + TL::TX - 1) / TL::TX;
    a.nby = (h->g.ny + TL::TY - 1) / TL::TY;
    a.nblocks = a.nbx * a.nby * ((h->g.nz + a.kchunk - 1) / a.kchunk);
    tile_base[q] = tt;
    tt += a.nblocks;
    // estimated cost of each tile: its planes plus the four warm-up planes,
    // times the waves that hold an active cell plus about three for the
    // cleaned halo ring
    for (int t = 0; t < a.nblocks; ++t) {
      const int by = (t / a.nbx) % a.nby, bz = t / (a.nbx * a.nby);
      const int rows = std::min(TL::TY, h->g.ny - by * TL::TY);
      const int planes = std::min(a.kchunk, h->g.nz - bz * a.kchunk);
      cost.push_back((double)(planes + 4) * ((rows * TL::TX + 63) / 64 + 3));
    }
    // per-instance lists padded to whole blocks of the ghost passes, so
    // that a block never spans two instances
    auto pad = [](int64_t n) { return (n + BATCH_TB - 1) / BATCH_TB * BATCH_TB; };
    info[q].nbuild = 0;
    if (!wrap) {
      // a WRAP_MIXED instance stores only its z-ghost planes (the first
      // cells in shell_cell's order) and cleans only its edge rings
      const int64_t nb = mixed ? (int64_t)2 * NG * (h->g.nx + 2 * NG) * (h->g.ny + 2 * NG) : shell_count(h);
      info[q].nbuild = nb;
      shell_base.push_back(ss);
      shell_inst.push_back(q);
      ss += pad(nb);
      if (mixed) {
        edge_base.push_back(ee);
        edge_inst.push_back(q);
        ee += pad((int64_t)(2 * (h->g.nx + 2) + 2 * h->g.ny) * (h->g.nz > 1 ? 2 : 1));
      } else {
        clean_base.push_back(cl);
        clean_inst.push_back(q);
        cl += pad(nb);
      }
    }
    if (h->bc_kind == TFF_BC_LOOP && h->bc_host[8] == 2.0) {
      floor_base.push_back(ff);
      floor_inst.push_back(q);
      ff += pad((int64_t)h->g.nx * h->g.ny);
    }
    for (int w = 0; w < 4; ++w) {
      BatchStage<S, P>& x = st[(size_t)w * n + q];
      x.a = a;
      x.ge_out = base_m<S>(h->ge);
      const bool first = w == 0;
      if (w <= 1) {
        set_coef<PR>(h, x.a, 1.0, 0.0, 1.0);
        x.a.ye = base_c<S>(h->es);
        x.a.yb = base_c<S>(h->bs);
        x.a.ge = first ? base_c<S>(h->es) : base_c<S>(h->ge);
        x.a.stage1 = 1;
        x.a.clean_in = first ? 0 : 1;
        x.a.oe = base_m<S>(h->w1e);
        x.a.ob = base_m<S>(h->w1b);
        // stages 2 and 3 read U^n E from ce in every step, so the first step
        // (whose input is already clean) copies it there as well
        x.a.u0e = base_m<S>(h->ce);
        x.ge_in = first ? base_c<S>(h->es) : base_c<S>(h->ge);
        x.prev_b = base_c<S>(h->bs);
      } else {
        const Vec& u0e = h->ce;  // steady state: stage 1 wrote the cleaned U^n E there
        x.a.stage1 = 0;
        x.a.clean_in = 1;
        x.a.ge = base_c<S>(h->ge);
        x.a.ue = base_c<S>(u0e);
        x.a.ub = base_c<S>(h->bs);
        x.a.u0e = nullptr;
        x.ge_in = base_c<S>(h->ge);
        if (w == 2) {
          set_coef<PR>(h, x.a, 3.0 / 4.0, 1.0 / 4.0, 1.0 / 4.0);
          x.a.ye = base_c<S>(h->w1e);
          x.a.yb = base_c<S>(h->w1b);
          x.a.oe = base_m<S>(h->w2e);
          x.a.ob = base_m<S>(h->w2b);
          x.prev_b = base_c<S>(h->w1b);
        } else {
          set_coef<PR>(h, x.a, 1.0 / 3.0, 2.0 / 3.0, 2.0 / 3.0);
          x.a.ye = base_c<S>(h->w2e);
          x.a.yb = base_c<S>(h->w2b);
          x.a.oe = base_m<S>(h->es);
          x.a.ob = base_m<S>(h->bs);
          x.prev_b = base_c<S>(h->w2b);
        }
      }
      // a WRAP_MIXED instance's carried z-ghost E is the stage input's raw
      // E (nothing cleans those planes)
      if (mixed) x.ge_in = x.a.ye;
    }
    info[q].per[0] = h->cfg.boundary[0] == TFF_PERIODIC;
    info[q].per[1] = h->cfg.boundary[1] == TFF_PERIODIC;
    info[q].per[2] = h->cfg.boundary[2] == TFF_PERIODIC;
    info[q].bc_kind = h->bc_kind;
    info[q].bcp = h->d_bc;
  }
  b->total_tiles = tt;
  {
    // segment x ends at the first item whose running cost reaches (x+1)/8
    // of the total
    double tot = 0.0;
    for (double c : cost) tot += c;
    double run = 0.0;
    int x = 0;
    b->qseg.seg[0] = 0;
    for (int64_t t = 0; t < tt && x < 7; ++t) {
      run += cost[(size_t)t];
      while (x < 7 && run >= tot * (x + 1) / 8.0) b->qseg.seg[++x] = t + 1;
    }
    while (x < 8) b->qseg.seg[++x] = tt;
    b->qseg.seg[8] = tt;
  }
  b->total_shell = ss;
  b->nghost = (int)shell_base.size();
  b->nclean = (int)clean_base.size();
  b->total_clean = cl;
  b->nedge = (int)edge_base.size();
  b->total_edge = ee;
  b->nfloor = (int)floor_base.size();
  b->total_floor = ff;
  // the batch stream is non-blocking: the old lists may only be freed, and
  // the new ones filled, in its order
  BHIP(b, hipStreamSynchronize(b->stream));
  hipFree(b->d_stage);
  hipFree(b->d_tile_base);
  hipFree(b->d_shell_base);
  hipFree(b->d_shell_inst);
  hipFree(b->d_clean_base);
  hipFree(b->d_clean_inst);
  hipFree(b->d_edge_base);
  hipFree(b->d_edge_inst);
  hipFree(b->d_floor_base);
  hipFree(b->d_floor_inst);
  hipFree(b->d_info);
  hipFree(b->d_fixmask);
  b->d_clean_base = b->d_edge_base = nullptr;
  b->d_clean_inst = b->d_edge_inst = nullptr;
  b->d_fixmask = nullptr;
  b->d_floor_base = nullptr;
  b->d_floor_inst = nullptr;
  b->d_stage = b->d_tile_base = nullptr;
  b->d_shell_base = nullptr;
  b->d_shell_inst = nullptr;
  b->d_info = nullptr;
  if (PR::null_fp64) {
    BHIP(b, hipMalloc(&b->d_fixmask, (size_t)tt * sizeof(unsigned long long)));
    for (int w = 0; w < 4; ++w)
      for (int q = 0; q < n; ++q)
        st[(size_t)w * n + q].a.fixmask = (unsigned long long*)b->d_fixmask + tile_base[q];
  }
  if (!b->d_queue) BHIP(b, hipMalloc((void**)&b->d_queue, BATCH_QWORDS * sizeof(unsigned)));
  BHIP(b, hipMemsetAsync(b->d_queue, 0, BATCH_QWORDS * sizeof(unsigned), b->stream));
  BHIP(b, hipMalloc(&b->d_stage, st.size() * sizeof(st[0])));
  BHIP(b, hipMemcpyAsync(b->d_stage, st.data(), st.size() * sizeof(st[0]), hipMemcpyHostToDevice, b->stream));
  BHIP(b, hipMalloc((void**)&b->d_tile_base, n * sizeof(int64_t)));
  BHIP(b, hipMemcpyAsync(b->d_tile_base, tile_base.data(), n * sizeof(int64_t), hipMemcpyHostToDevice, b->stream));
  BHIP(b, hipMalloc(&b->d_info, n * sizeof(BatchInfo)));
  BHIP(b, hipMemcpyAsync(b->d_info, info.data(), n * sizeof(BatchInfo), hipMemcpyHostToDevice, b->stream));
  if (b->nghost) {
    BHIP(b, hipMalloc((void**)&b->d_shell_base, shell_base.size() * sizeof(int64_t)));
    BHIP(b, hipMemcpyAsync(b->d_shell_base, shell_base.data(), shell_base.size() * sizeof(int64_t), hipMemcpyHostToDevice, b->stream));
    BHIP(b, hipMalloc((void**)&b->d_shell_inst, shell_inst.size() * sizeof(int)));
    BHIP(b, hipMemcpyAsync(b->d_shell_inst, shell_inst.data(), shell_inst.size() * sizeof(int), hipMemcpyHostToDevice, b->stream));
  }
  if (b->nclean) {
    BHIP(b, hipMalloc((void**)&b->d_clean_base, clean_base.size() * sizeof(int64_t)));
    BHIP(b, hipMemcpyAsync(b->d_clean_base, clean_base.data(), clean_base.size() * sizeof(int64_t),
                           hipMemcpyHostToDevice, b->stream));
    BHIP(b, hipMalloc((void**)&b->d_clean_inst, clean_inst.size() * sizeof(int)));
    BHIP(b, hipMemcpyAsync(b->d_clean_inst, clean_inst.data(), clean_inst.size() * sizeof(int),
                           hipMemcpyHostToDevice, b->stream));
  }
  if (b->nedge) {
    BHIP(b, hipMalloc((void**)&b->d_edge_base, edge_base.size() * sizeof(int64_t)));
    BHIP(b, hipMemcpyAsync(b->d_edge_base, edge_base.data(), edge_base.size() * sizeof(int64_t),
                           hipMemcpyHostToDevice, b->stream));
    BHIP(b, hipMalloc((void**)&b->d_edge_inst, edge_inst.size() * sizeof(int)));
    BHIP(b, hipMemcpyAsync(b->d_edge_inst, edge_inst.data(), edge_inst.size() * sizeof(int),
                           hipMemcpyHostToDevice, b->stream));
  }
  if (b->nfloor) {
    BHIP(b, hipMalloc((void**)&b->d_floor_base, floor_base.size() * sizeof(int64_t)));
    BHIP(b, hipMemcpyAsync(b->d_floor_base, floor_base.data(), floor_base.size() * sizeof(int64_t),
                           hipMemcpyHostToDevice, b->stream));
    BHIP(b, hipMalloc((void**)&b->d_floor_inst, floor_inst.size() * sizeof(int)));
    BHIP(b, hipMemcpyAsync(b->d_floor_inst, floor_inst.data(), floor_inst.size() * sizeof(int),
                           hipMemcpyHostToDevice, b->stream));
  }
  BHIP(b, hipStreamSynchronize(b->stream));  // the host vectors are in use until here
  return batch_grid<PR>(b);
}

template <class PR>
int batch_build(tff_batch* b) {
  return with_tile<int>(b->tile, [&](auto tl) { return batch_build_t<PR, decltype(tl)>(b); });
}

template <class PR>
int batch_ghosts(tff_batch* b, int which) {
  using S = typename PR::S;
  using P = typename PR::P;
  if (!b->nghost) return TFF_OK;
  const int n = (int)b->inst.size();
  const BatchStage<S, P>* st = (const BatchStage<S, P>*)b->d_stage + (size_t)which * n;
  const int tb = BATCH_TB;
  if (b->nfloor)
    hipLaunchKernelGGL(k_batch_floor<PR>, dim3((unsigned)((b->total_floor + tb - 1) / tb)), dim3(tb), 0, b->stream,
                       st, (const BatchInfo*)b->d_info, b->d_floor_base, b->d_floor_inst, b->nfloor,
                       b->total_floor, b->d_step);
  const unsigned nb = (unsigned)((b->total_shell + tb - 1) / tb);
  hipLaunchKernelGGL(k_batch_ghost_build<PR>, dim3(nb), dim3(tb), 0, b->stream, st, (const BatchInfo*)b->d_info,
                     b->d_shell_base, b->d_shell_inst, b->nghost, b->total_shell, b->d_step);
  if (b->nclean)
    hipLaunchKernelGGL(k_batch_ghost_clean<PR>, dim3((unsigned)((b->total_clean + tb - 1) / tb)), dim3(tb), 0,
                       b->stream, st, (const BatchInfo*)b->d_info, b->d_clean_base, b->d_clean_inst, b->nclean,
                       b->total_clean);
  if (b->nedge)
    hipLaunchKernelGGL(k_batch_edge_clean<PR>, dim3((unsigned)((b->total_edge + tb - 1) / tb)), dim3(tb), 0,
                       b->stream, st, b->d_edge_base, b->d_edge_inst, b->nedge, b->total_edge);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return bfail(b, TFF_EHIP, std::string("batch ghost passes: ") + hipGetErrorString(e));
  return TFF_OK;
}

// one timestep of every instance
template <class PR>
int batch_enqueue_step(tff_batch* b) {
  int rc;
  const int s1 = b->first ? 0 : 1;
  if ((rc = batch_launch_stage<PR>(b, s1))) return rc;
  if ((rc = batch_ghosts<PR>(b, s1))) return rc;
  for (int w = 2; w <= 3; ++w) {
    if ((rc = batch_launch_stage<PR>(b, w))) return rc;
    if ((rc = batch_ghosts<PR>(b, w))) return rc;
  }
  hipLaunchKernelGGL(k_bump_step, dim3(1), dim3(64), 0, b->stream, b->d_step, 1);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return bfail(b, TFF_EHIP, hipGetErrorString(e));
  return TFF_OK;
}

// Autotune the batch: every tile shape and z-chunk on the stage-1 launch of
// a running step (reads the state, writes the stage-1 output and the cleaned
// U^n E, which the clean state does not use: no state is touched).
template <class PR>
int batch_autotune(tff_batch* b) {
  if (b->tile >= 0) return batch_build<PR>(b);
  const int kcs[] = {8, 16, 32, 64};
  float best = 1e30f;
  int bt = -1, bk = 0;
  hipEvent_t e0, e1;
  BHIP(b, hipEventCreate(&e0));
  BHIP(b, hipEventCreate(&e1));
  for (int t = 0; t < NTILES; ++t) {
    if (!batch_tile_lds_ok<PR>(b, t)) continue;
    for (int kc : kcs) {
      b->tile = t;
      b->kchunk = kc;
      int rc = batch_build<PR>(b);
      if (rc) return rc;
      if ((rc = batch_launch_stage<PR>(b, 1))) return rc;
      float ms = 1e30f;
      for (int round = 0; round < 2; ++round) {
        BHIP(b, hipEventRecord(e0, b->stream));
        if ((rc = batch_launch_stage<PR>(b, 1))) return rc;
        BHIP(b, hipEventRecord(e1, b->stream));
        BHIP(b, hipEventSynchronize(e1));
        float m = 0.f;
        BHIP(b, hipEventElapsedTime(&m, e0, e1));
        ms = std::min(ms, m);
      }
      if (ms < best) {
        best = ms;
        bt = t;
        bk = kc;
      }
    }
  }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  if (bt < 0) return bfail(b, TFF_EINVAL, "no tile shape fits the device's LDS");
  b->tile = bt;
  b->kchunk = bk;
  return batch_build<PR>(b);
}

template <class PR>
int batch_run_t(tff_batch* b, int first_step, int nsteps) {
  int rc;
  // every run starts from the cleaned (reference) state of each instance, so
  // all instances share the first-step descriptors; the buffer roles this
  // leaves are what the descriptors and the captured graph refer to
  for (tff_handle* h : b->inst) {
    BHIP(b, hipStreamSynchronize(h->stream));
    if ((rc = ensure_t0<PR>(h))) return bfail(b, rc, h->err);
    if ((rc = ensure_clean_t<PR>(h))) return bfail(b, rc, h->err);
    BHIP(b, hipStreamSynchronize(h->stream));
  }
  b->first = true;
  if (b->graph) {
    hipGraphExecDestroy(b->graph);
    b->graph = nullptr;
  }
  if (b->tile < 0) {
    if ((rc = batch_autotune<PR>(b))) return rc;
  } else if ((rc = batch_build<PR>(b))) {
    return rc;
  }
  hipLaunchKernelGGL(k_set_int, dim3(1), dim3(64), 0, b->stream, b->d_step, first_step);
  for (int s = 0; s < nsteps; ++s) {
    if (b->first) {
      if ((rc = batch_enqueue_step<PR>(b))) return rc;
      b->first = false;
      continue;
    }
    if (!b->graph) {
      hipGraph_t g;
      BHIP(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
      rc = batch_enqueue_step<PR>(b);
      hipError_t e = hipStreamEndCapture(b->stream, &g);
      if (rc) return rc;
      if (e != hipSuccess) return bfail(b, TFF_EHIP, "hipStreamEndCapture failed");
      e = hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0);
      hipGraphDestroy(g);
      if (e != hipSuccess) return bfail(b, TFF_EHIP, "hipGraphInstantiate failed");
    }
    BHIP(b, hipGraphLaunch(b->graph, b->stream));
  }
  BHIP(b, hipStreamSynchronize(b->stream));
  for (tff_handle* h : b->inst) {
    h->state_clean = false;
    h->tile = b->tile;
    h->kchunk = b->kchunk;
    invalidate_graphs(h);
  }
  return TFF_OK;
}

template <class PR>
struct DoBatchRun {
  static int run(tff_batch* b, int first, int n) { return batch_run_t<PR>(b, first, n); }
};

template <template <class> class F, class... A>
int bdispatch(tff_batch* b, A... a) {
  switch (b->precision) {
    case TFF_FP64: return F<PrecF64>::run(b, a...);
    case TFF_FP32: return F<PrecF32>::run(b, a...);
    default: return F<PrecF32Bulk>::run(b, a...);
  }
}

}  // namespace

extern "C" {

int tff_batch_create(const tff_instance* in, int n, int precision, int device, tff_batch** out) {
  if (!in || n < 1 || !out) return TFF_EINVAL;
  *out = nullptr;
  if (precision == TFF_FP32_COMP) return TFF_EINVAL;  // the batched kernels have no lo words
  tff_batch* b = new tff_batch();
  b->precision = precision;
  b->device = device;
  int rc = TFF_OK;
  for (int q = 0; q < n && !rc; ++q) {
    if (in[q].bc_kind != TFF_BC_NONE && in[q].bc_kind != TFF_BC_LOOP) {
      rc = TFF_EINVAL;
      break;
    }
    if (in[q].bc_kind == TFF_BC_LOOP && (in[q].nparams < 9 || in[q].params[8] < 0.0 || in[q].params[8] > 2.0)) {
      rc = TFF_EINVAL;
      break;
    }
    for (int a = 0; a < 3; ++a)
      if (in[q].boundary[a] == TFF_ABSORB) rc = TFF_EINVAL;
    if (rc) break;
    tff_config c{};
    c.nx = in[q].nx;
    c.ny = in[q].ny;
    c.nz = in[q].nz;
    for (int a = 0; a < 3; ++a) c.boundary[a] = in[q].boundary[a];
    c.absorb_buff = 0;
    c.absorb_fixed = 0;
    c.precision = precision;
    c.device = device;
    c.cc = in[q].cc;
    tff_handle* h = nullptr;
    if ((rc = tff_create(&c, &h))) break;
    b->inst.push_back(h);
    if ((rc = tff_set_bc(h, in[q].bc_kind, in[q].params, in[q].nparams))) break;
  }
  if (!rc && hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) rc = TFF_EHIP;
  if (!rc && hipMalloc((void**)&b->d_step, sizeof(int)) != hipSuccess) rc = TFF_ENOMEM;
  if (rc) {
    tff_batch_destroy(b);
    return rc;
  }
  *out = b;
  return TFF_OK;
}

void tff_batch_destroy(tff_batch* b) {
  if (!b) return;
  if (b->stream) hipStreamSynchronize(b->stream);
  if (b->graph) hipGraphExecDestroy(b->graph);
  for (tff_handle* h : b->inst) tff_destroy(h);
  hipFree(b->d_tile_base);
  hipFree(b->d_shell_base);
  hipFree(b->d_shell_inst);
  hipFree(b->d_clean_base);
  hipFree(b->d_clean_inst);
  hipFree(b->d_edge_base);
  hipFree(b->d_edge_inst);
  hipFree(b->d_floor_base);
  hipFree(b->d_floor_inst);
  hipFree(b->d_info);
  hipFree(b->d_stage);
  hipFree(b->d_fixmask);
  hipFree(b->d_queue);
  hipFree(b->d_step);
  if (b->stream) hipStreamDestroy(b->stream);
  delete b;
}

const char* tff_batch_last_error(const tff_batch* b) { return b ? b->err.c_str() : "null batch"; }

tff_handle* tff_batch_instance(tff_batch* b, int i) {
  if (!b || i < 0 || i >= (int)b->inst.size()) return nullptr;
  hipStreamSynchronize(b->stream);
  return b->inst[i];
}

int tff_batch_run(tff_batch* b, int first_step, int nsteps) {
  if (!b || nsteps < 0) return TFF_EINVAL;
  if (nsteps == 0) return TFF_OK;
  return bdispatch<DoBatchRun>(b, first_step, nsteps);
}

int tff_batch_get_tuning(tff_batch* b, int* tile, int* tx, int* ty, int* kchunk) {
  if (!b || !tile || !tx || !ty || !kchunk) return TFF_EINVAL;
  *tile = b->tile;
  *tx = b->tile >= 0 ? TILE_TX[b->tile] : 0;
  *ty = b->tile >= 0 ? TILE_TY[b->tile] : 0;
  *kchunk = b->kchunk;
  return TFF_OK;
}

}  // extern "C"
