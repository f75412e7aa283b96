/* tff_io.h -- the output files of tristan-ff's writeOutput
 * (src/io/write_output.F90:27-43) for one meshblock, through the serial
 * HDF5 C API: params.NNNNN, flds.tot.NNNNN with its XDMF wrapper, and
 * domain.NNNNN, NNNNN being the output index (five digits, zero padded).
 *
 * The reference writes with parallel HDF5 from every rank; on one GPU the
 * whole box is one meshblock, so one process writes each file whole.  The
 * dataset names, types, shapes and the XDMF text are the reference's.
 *
 * Conventions as in tff.h: TFF_OK (0) or a negative TFF_E* code;
 * tff_io_last_error() gives the text of the last failure on this thread. */
#ifndef TFF_IO_H
#define TFF_IO_H

#include "tff.h"

#ifdef __cplusplus
extern "C" {
#endif

/* What var1..var3 of flds.tot hold, i.e. which user file's userOutput
 * (write_output.F90:375-383) the caller's problem has. */
#define TFF_USEROUT_NONE 0    /* all zero (user_emwave and user_dummy have no userOutput) */
#define TFF_USEROUT_RHO 1     /* var1 = rho, var2 = var3 = 0 (user_loop.F90:168-180, user_psr.F90:386-398) */
#define TFF_USEROUT_UNIFORM 2 /* var1 = rho, var2 = |E x B|_perp / (r |B|^2), var3 = 0 (user_uniform.F90:151-173) */

const char* tff_io_last_error(void);

/* writeParams_hdf5 (write_output.F90:60-114): `dir`/params.NNNNN holding
 * the integer dataset "timestep" and one dataset per deck parameter, named
 * "group:name" with the group abbreviated as the reference's registry does
 * (cpu, grd, tim, alg, out, prb, ...).  `names` holds n NUL-terminated
 * strings one after another; types[q] is 1 (integer, ivals[q]), 2 (real,
 * written as FP32 from rvals[q]) or 3 (logical, written as an integer from
 * ivals[q]).  Writes nothing but the file. */
int tff_io_write_params(const char* dir, int index, int timestep, int n, const char* names, const int* types,
                        const int* ivals, const float* rvals);

/* writeFields_hdf5 and writeXDMF_hdf5 (write_output.F90:116-429) for a
 * meshblock at the origin: `dir`/flds.tot.NNNNN with the FP32 datasets ex
 * ey ez bx by bz var1 var2 var3 xx yy zz, each of dims[0] x dims[1] x
 * dims[2] nodes (x fastest; the Fortran shape (ni, nj, nk), which HDF5's C
 * view lists as (nk, nj, ni)), and `dir`/flds.tot.NNNNN.xdmf.  data[0..8]
 * are ex..bz and var1..var3; xx yy zz are the nodes' cell indices
 * (i1, j1, k1) * istep. */
int tff_io_write_fields(const char* dir, int index, int istep, const int dims[3], const float* const data[9]);

/* writeDomain_hdf5 (write_output.F90:431-500): `dir`/domain.NNNNN with the
 * integer datasets x0 y0 z0 sx sy sz, one entry per meshblock; blocks holds
 * (x0, y0, z0, sx, sy, sz) of each of the n meshblocks. */
int tff_io_write_domain(const char* dir, int index, int n, const int* blocks);

/* writeOutput(time) for a handle: the three files above, the field
 * datasets from tff_output_fields, var1..var3 composed from
 * tff_output_samples as user_output says, the handle's box as the one
 * meshblock.  istep >= 2 (the reference's istep = 1 debug layout, which
 * includes ghosts, is rejected).  The output directory is created if
 * missing (initialize.F90:242-254 does so at start-up). */
int tff_io_write_output(tff_handle* h, const char* dir, int index, int timestep, int istep, int user_output,
                        int n, const char* names, const int* types, const int* ivals, const float* rvals);

#ifdef __cplusplus
}
#endif
#endif
