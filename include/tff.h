/*
 * tff.h -- C ABI of the MI355X force-free step.
 *
 * The reference (tristan-ff) exposes its hot path as Fortran module
 * procedures acting on module-global arrays; it has no C interface.  Each
 * entry point below replaces one of those procedures and keeps its meaning,
 * so that a Fortran host binds them through ISO_C_BINDING (see
 * tristan-ff_amd/fortran/m_tff.F90) and keeps the reference's call order.
 *
 * Conventions
 *  - Every function returns TFF_OK (0) or a negative TFF_E* code; nothing
 *    aborts.  tff_last_error() gives the text of the last failure.
 *  - A handle owns all device memory of one meshblock and one HIP stream.
 *    Calls are asynchronous on that stream except tff_upload, tff_download,
 *    tff_output_fields, tff_reduce and tff_sync, which return after the
 *    stream has drained.
 *  - A handle is not thread-safe: call it from one host thread.
 *  - Host arrays passed to tff_upload / tff_download hold one field
 *    component in the reference's layout: Fortran order (x fastest),
 *    indices -2..n-1+2 on every axis, i.e. (nx+4)*(ny+4)*(nz+4) values
 *    starting at element (-2,-2,-2) (src/objects/fields.F90:14-27,
 *    src/logistics/initialize.F90:192-199; NGHOST=2, src/CMakeLists.txt:19).
 */
#ifndef TFF_H
#define TFF_H

#ifdef __cplusplus
extern "C" {
#endif

#define TFF_NGHOST 2

/* return codes */
#define TFF_OK 0
#define TFF_EINVAL -1   /* bad argument or configuration */
#define TFF_EHIP -2     /* a HIP runtime call failed */
#define TFF_ENOMEM -3   /* device allocation failed */
#define TFF_ESTATE -4   /* call not valid in the handle's current state */

/* precision of the path (the reference computes in default REAL = FP32,
 * CMakeLists.txt:23,25) */
#define TFF_FP64 0      /* FP64 storage and arithmetic */
#define TFF_FP32 1      /* FP32 everywhere, operation order of the reference */
#define TFF_FP32_BULK 2 /* FP32 storage and arithmetic with fused
                           multiply-adds; the stage update in the form
                           U^n + c ((X - U^n) + dU), exact in scale; cells
                           near a magnetic null redo their projection and
                           currents in FP64; FP64 reductions */
#define TFF_FP32_COMP 3 /* FP32-bulk stencils on a compensated state: each
                           field value is a pair of FP32 words hi + lo, the
                           stencils read hi, the stage accumulation adds the
                           increment to the pair by two-sum and the E.B
                           projection works on hi + lo in FP64; twice the
                           storage; the per-procedure kernels only (no fused
                           path, no batches).  An FP64 upload splits each
                           value into the pair, an FP64 download returns
                           hi + lo.  Cells a boundary hook writes keep the
                           hook's FP32 value with the lo they had. */

/* boundary flags, src/objects/domain.F90:31-33 and initialize.F90:139-148 */
#define TFF_OPEN 0
#define TFF_PERIODIC 1
#define TFF_ABSORB 2

/* field component ids for upload/download (src/objects/fields.F90:14-27) */
#define TFF_EX 0
#define TFF_EY 1
#define TFF_EZ 2
#define TFF_BX 3
#define TFF_BY 4
#define TFF_BZ 5
#define TFF_B0X 6
#define TFF_B0Y 7
#define TFF_B0Z 8
#define TFF_ENX 9
#define TFF_ENY 10
#define TFF_ENZ 11
#define TFF_BNX 12
#define TFF_BNY 13
#define TFF_BNZ 14

/* host element types */
#define TFF_HOST_F64 0
#define TFF_HOST_F32 1

/* boundary-condition kernels (userFieldBoundaryConditions of the user files) */
#define TFF_BC_NONE 0    /* user_emwave.F90:75-78, user_dummy.F90:51-54 */
#define TFF_BC_LOOP 1    /* user_loop.F90:101-164 */
#define TFF_BC_PSR 2     /* user_psr.F90:72-332 */
#define TFF_BC_UNIFORM 3 /* user_uniform.F90:77-147 */

typedef struct tff_handle tff_handle;

typedef struct tff_config {
  int nx, ny, nz;      /* interior cells: this_meshblock%ptr%{sx,sy,sz} */
  int boundary[3];     /* boundary_x/y/z: TFF_OPEN, TFF_PERIODIC, TFF_ABSORB;
                          any TFF_ABSORB makes all three absorbing as
                          initialize.F90:143-148 does */
  int absorb_buff;     /* <grid> absorb_buff (default 20, initialize.F90:147) */
  int absorb_fixed;    /* 0: as coded -- rk3Update's absorbing layer uses i
                          for jg (field_solver.F90:211) and an unclamped
                          delta, and cleanEpar divides by |B|^2 with no
                          floor (:300); 1: fixed -- jg from j, delta clamped
                          to 1 (fields beyond r_max are zeroed) and cleanEpar
                          skips the projection where |B|^2 is exactly 0 */
  int precision;       /* TFF_FP64, TFF_FP32, TFF_FP32_BULK, TFF_FP32_COMP */
  int device;          /* HIP device ordinal */
  double cc;           /* <algorithm> c, the CFL number CC (global_namespace.F90:14) */
} tff_config;

/* Allocate a meshblock on the device.  Replaces the field part of
 * initializeFields (initialize.F90:201-240); all arrays start at zero. */
int tff_create(const tff_config* cfg, tff_handle** out);
void tff_destroy(tff_handle* h);
const char* tff_last_error(const tff_handle* h);

/* Copy one component between a host array (layout above) and the device.
 * Replaces direct access to the m_fields module arrays by userInitFields
 * (e.g. user_psr.F90:39-69) and by the output writer. */
int tff_upload(tff_handle* h, int field, const void* host, int host_dtype);
int tff_download(tff_handle* h, int field, void* host, int host_dtype);

/* Register the problem's boundary kernel and its parameters (the values the
 * user file's userReadInput computes), in this order:
 *  TFF_BC_NONE     none
 *  TFF_BC_LOOP     xc, yc1, yc2, zc, bnorm, r0, beta_twist, t0, mode
 *                  (user_loop.F90:27-42 and :101-164).  mode 0: the
 *                  floor/top copy as shipped; 1: the footpoint driver of the
 *                  commented block :122-150 on the floor ghosts; 2: that
 *                  driver's E on z <= 0 (floor plane included, its bz left
 *                  to the update so div B is kept) with a conducting top,
 *                  the boundary form of user_uniform.F90:113-142
 *  TFF_BC_PSR      psr_radius, psr_angle [rad], psr_omega, xc, yc, zc
 *                  (user_psr.F90:26-37 and :72-332)
 *  TFF_BC_UNIFORM  xc, yc, r0, beta_twist, t0, bnorm
 *                  (user_uniform.F90:27-69 and :77-147)
 * Positions are in cells of this meshblock (global = local here). */
int tff_set_bc(tff_handle* h, int kind, const double* params, int nparams);

/* field_solver.F90:15-33 initRKstep: En,Bn <- E,B. */
int tff_init_rk_step(tff_handle* h);
/* field_solver.F90:35-42 rk3Step(rk_c1, rk_c2, rk_c3): computeRho, rk3Push
 * and rk3Update in one pass.  U <- c1*Un + c2*U + c3*L(U) on the interior. */
int tff_rk3_step(tff_handle* h, double c1, double c2, double c3);
/* field_solver.F90:273-335 cleanEpar: remove E parallel to B (interior). */
int tff_clean_epar(tff_handle* h);
/* exchange_fields.F90:11-197 exchangeFields([exchangeB0]) on one meshblock:
 * ghost layers of periodic axes take the wrapped interior; ghosts of
 * non-periodic axes are left as they are. */
int tff_exchange_fields(tff_handle* h, int exchange_b0);
/* userFieldBoundaryConditions(step) of the registered problem. */
int tff_user_bc(tff_handle* h, int step);

/* One whole timestep in the order of tristanmainloop.F90:50-83 (three
 * stages, each followed by BC, ghost fill, cleanEpar, ghost fill). */
int tff_step(tff_handle* h, int timestep);
/* nsteps timesteps first_step, first_step+1, ... replayed from a captured
 * HIP graph. */
int tff_run(tff_handle* h, int first_step, int nsteps);

/* The field datasets of writeOutput's flds.tot file for one meshblock at
 * the origin: writeFields_hdf5 (src/io/write_output.F90:185-429) with
 * output_istep >= 2, i.e. at nodes 0, istep, 2 istep, ... of each axis, E
 * interpolated from the edges and B from the faces at offset 0
 * (interpFromEdges / interpFromFaces, src/tools/helpers.F90:93-257), in the
 * reference's operation order and the handle's storage precision.
 * tff_output_dims gives the node counts ceil(n / istep) per axis.
 * tff_output_fields fills out[0..5] (ex ey ez bx by bz), each
 * dims[0]*dims[1]*dims[2] values of host_dtype, x fastest (sm_arr(i1, j1,
 * k1)); it returns after the copies.  The HDF5 file itself (and the var1-3
 * user datasets and xx/yy/zz coordinates) stay with the host.
 * output_istep = 1 (the reference's ghost-including debug layout) is
 * rejected with TFF_EINVAL. */
int tff_output_dims(tff_handle* h, int istep, int dims[3]);
int tff_output_fields(tff_handle* h, int istep, void* const out[6], int host_dtype);

/* What userOutput(var, ...) of the user files reads at the same nodes
 * (write_output.F90:375-383): out[0..5] the raw ex ey ez bx by bz at cell
 * (i, j, k) = (i1, j1, k1) * istep, out[6] the charge density there in
 * computeRho's form (field_solver.F90:57-59) of the state written.  The
 * reference's userOutput reads the rho its last rk3Push left (the third
 * stage's input); this is the rho of the output state, the value of the
 * call to computeRho that tristanmainloop.F90:92 leaves commented out.
 * Same layout and dtypes as tff_output_fields. */
int tff_output_samples(tff_handle* h, int istep, void* const out[7], int host_dtype);

/* Device staging of tff_output_fields_noalloc: 24 MiB, i.e. six datasets
 * of up to 1,048,576 nodes as FP32 or 524,288 as FP64. */
#define TFF_NOALLOC_SCRATCH_BYTES (24 * 1024 * 1024)
/* tff_output_fields without any device allocation: no counterpart in the
 * reference.  It is for callers that cannot allocate device memory (tight
 * output loops, harnesses that forbid hipMalloc).  Same arguments, results
 * and errors as tff_output_fields; it stages through a fixed device array
 * of TFF_NOALLOC_SCRATCH_BYTES that each GPU holds until the process ends,
 * and rejects outputs larger than that with TFF_EINVAL.  Every call on one
 * GPU shares that array: calls that overlap on a GPU (other threads or
 * handles) race. */
int tff_output_fields_noalloc(tff_handle* h, int istep, void* const out[6], int host_dtype);

/* Initial conditions set in HBM, as the user files' userInitFields set
 * them, without host arrays: the state E, B and the background B0 of the
 * whole box (ghosts included), each value computed in FP64 and stored in
 * the handle's precision.  The state is then clean, as after tff_upload.
 *  TFF_IC_LOOP     user_loop.F90:44-93, params xc, yc1, yc2, zc, bnorm:
 *                  the two-source potential field, E = 0, B0 = B;
 *  TFF_IC_PSR      user_psr.F90:39-69, params psr_radius, psr_angle [rad],
 *                  psr_omega, xc, yc, zc: the dipole on the interior
 *                  cells, zero ghosts, E = 0, B0 = 0;
 *  TFF_IC_UNIFORM  user_uniform.F90:43-69, params xc, yc, r0, bnorm:
 *                  bz = bnorm / (1 + r/r0), E = 0, B0 = B. */
#define TFF_IC_LOOP 1
#define TFF_IC_PSR 2
#define TFF_IC_UNIFORM 3
int tff_init_fields(tff_handle* h, int kind, const double* params, int nparams);
/* A periodic field given as Fourier modes (no reference counterpart: the
 * Alfvenic perturbation of the measured workloads): component q of
 * interior cell (i, j, k) = background[q] + sum over modes of
 * Re[c_q exp(2 pi i (kx i/nx + ky j/ny + kz k/nz))], q = ex ey ez bx by bz;
 * ghosts zero; B0 = background[3..5] everywhere.  Each mode is 15 doubles:
 * kx, ky, kz (integers, |k| (n + 4) < 2^31, else TFF_EINVAL), then (re, im)
 * of c for the six components. */
int tff_init_spectrum(tff_handle* h, const double* modes, int nmodes, const double background[6]);

/* FP64 diagnostics over the interior (no reference counterpart):
 * out[0] = 1/2 sum E^2, out[1] = 1/2 sum B^2, out[2] = max |div_h B|,
 * out[3] = sqrt(mean (div_h B)^2). */
int tff_reduce(tff_handle* h, double out[4]);

/* Wait for all work queued on the handle's stream. */
int tff_sync(tff_handle* h);

/* Timing on the handle's stream (for benchmarks): record event slot
 * 0..15, then read the time between two recorded slots in milliseconds. */
int tff_event_record(tff_handle* h, int slot);
int tff_event_elapsed(tff_handle* h, int slot_a, int slot_b, float* ms);
/* When enabled, every launch of the stage kernel is bracketed by HIP events
 * on the handle's stream and tff_stage_kernel_time() returns the summed
 * kernel time and the number of launches since the last reset. */
int tff_set_kernel_timing(tff_handle* h, int on);
int tff_stage_kernel_time(tff_handle* h, double* total_ms, long* launches);

/* Path of tff_step / tff_run.  fused = 1 (default): one fused kernel per
 * stage applies the previous stage's cleanEpar on the fly, so the state
 * between steps is the raw output of stage 3 and is cleaned only when the
 * host reads it or calls a per-procedure entry point.  fused = 0: the
 * per-procedure kernels in the order of tristanmainloop.F90:50-83.  Both
 * give the same numbers bit for bit. */
int tff_set_path(tff_handle* h, int fused);
/* Tile shape (index into the built shapes; -1 = let the autotuner pick on
 * the next step) and z-chunk (a multiple of 4 in 4..64) of the fused stage
 * kernel, and the
 * current choice (tile index, its x and y extent, z-chunk). */
int tff_set_tuning(tff_handle* h, int tile, int kchunk);
int tff_get_tuning(tff_handle* h, int* tile, int* tx, int* ty, int* kchunk);
/* Phase timers for a timing report like the reference's makeReport
 * (tristanmainloop.F90:108-151): while on, every step runs eagerly (no
 * graph replay) with HIP events around its phases -- TFF_PHASE_SOLVER the
 * stage kernels (rk3Push, rk3Update, cleanEpar; fld_solver without the
 * exchange), TFF_PHASE_HOOK the user boundary kernels (usr_funcs),
 * TFF_PHASE_EXCHANGE the ghost exchange and ghost-cell passes.
 * tff_phase_times synchronizes, returns the GPU milliseconds of each phase
 * since the last call (or since timing was switched on) and restarts the
 * sums. */
#define TFF_PHASES 3
#define TFF_PHASE_SOLVER 0
#define TFF_PHASE_HOOK 1
#define TFF_PHASE_EXCHANGE 2
int tff_set_phase_timing(tff_handle* h, int on);
int tff_phase_times(tff_handle* h, double ms[3]);

/* Device layout of one field component (no reference counterpart: the
 * reference's arrays are host Fortran arrays): out[0] = padded x-row pitch,
 * out[1] = plane pitch, out[2] = offset of interior cell (0,0,0), out[3] =
 * elements per component (nalloc), all in elements; out[4] = bytes per
 * element. */
int tff_get_layout(tff_handle* h, long long out[5]);
/* The box's interior size n[0..2] (the deck's mx0, my0, mz0) and the
 * precision it was created with. */
int tff_get_size(tff_handle* h, int n[3], int* precision);

/* ---- batched path (parameter scans) ----------------------------------
 * Many independent meshblocks of uneven size advanced together: per RK
 * stage one persistent launch of the fused stage kernel works through a
 * device-side list of (instance, tile) items, and the ghost passes of all
 * instances run as two batched launches; steady-state steps replay one
 * captured HIP graph.  Each instance is an ordinary handle
 * (tff_batch_instance) for upload, download, reductions and the pre-loop
 * exchange + hook; supported hooks are TFF_BC_NONE and TFF_BC_LOOP on
 * periodic or open axes. */
typedef struct tff_batch tff_batch;
typedef struct tff_instance {
  int nx, ny, nz;
  int boundary[3];
  double cc;
  int bc_kind;
  int nparams;
  double params[16];
} tff_instance;
int tff_batch_create(const tff_instance* inst, int n, int precision, int device, tff_batch** out);
void tff_batch_destroy(tff_batch* b);
const char* tff_batch_last_error(const tff_batch* b);
/* The handle of instance i (owned by the batch; do not destroy it).  Waits
 * for the batch's queued work. */
tff_handle* tff_batch_instance(tff_batch* b, int i);
/* nsteps timesteps of every instance; returns when they are done. */
int tff_batch_run(tff_batch* b, int first_step, int nsteps);
int tff_batch_get_tuning(tff_batch* b, int* tile, int* tx, int* ty, int* kchunk);

#ifdef __cplusplus
}
#endif
#endif /* TFF_H */
