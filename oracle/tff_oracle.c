/*
 * tff_oracle.c -- CPU restatement of tristan-ff's force-free step.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the
 * checker.  The product path (tristan-ff_amd/csrc, libtff.so) never calls it.
 *
 * Every function restates one routine of the reference (paths relative to
 * the reference root) with the same loop bounds, the same operand order and
 * the same floating-point operation order, so that with -ffp-contract=off
 * the FP64 build is bit-for-bit what the reference computes under
 * -fdefault-real-8 -fdefault-double-8.  The library is compiled twice:
 *   ORC_REAL=double -> liboracle_f64.so  (the FP64 oracle)
 *   ORC_REAL=float  -> liboracle_f32.so  (the reference as shipped: default
 *                                         REAL is FP32, CMakeLists.txt:23,25)
 *
 * Array layout is the reference's (src/objects/fields.F90:14-27,
 * src/logistics/initialize.F90:192-199): column-major, x fastest, indices
 * -NGHOST..s-1+NGHOST on every axis, NGHOST = 2 (src/CMakeLists.txt:19).
 *
 * Parity status of each function (how it is pinned) is listed in the
 * header comment of oracle/oracle.py.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifndef ORC_REAL
#define ORC_REAL double
#endif
typedef ORC_REAL real;

#define NG 2

/* src/defs.F90:4 -- TINY is a default-REAL literal */
#define ORC_TINY ((real)1e-7)

enum {
  F_EX, F_EY, F_EZ, F_BX, F_BY, F_BZ,
  F_B0X, F_B0Y, F_B0Z,
  F_ENX, F_ENY, F_ENZ, F_BNX, F_BNY, F_BNZ,
  F_DEX, F_DEY, F_DEZ, F_DBX, F_DBY, F_DBZ,
  F_RHO, F_COUNT
};

/* boundary-condition kinds (user/*.F90 userFieldBoundaryConditions) */
enum { BC_NONE = 0, BC_LOOP = 1, BC_PSR = 2, BC_UNIFORM = 3 };

typedef struct {
  int nx, ny, nz;          /* this_meshblock%ptr%{sx,sy,sz} */
  int x0, y0, z0;          /* this_meshblock%ptr%{x0,y0,z0} */
  int gx, gy, gz;          /* global_mesh%{sx,sy,sz} */
  int bnd[3];              /* boundary_{x,y,z}: 0 open, 1 periodic, 2 absorb */
  int absorb_buff;
  int absorb_fixed;        /* 0: as coded (rk3Update quirks); 1: fixed */
  real cc, ccinv;
  int64_t sx, sy, sz, ncell;
  real *f[F_COUNT];
  int bc_kind;
  double bcp[16];
} orc_t;

#define IDX(o, i, j, k) \
  ((int64_t)((i) + NG) + (o)->sx * ((int64_t)((j) + NG) + (o)->sy * (int64_t)((k) + NG)))

orc_t *orc_create(int nx, int ny, int nz, int x0, int y0, int z0, int gx, int gy,
                  int gz, int bndx, int bndy, int bndz, int absorb_buff,
                  int absorb_fixed, double cc) {
  orc_t *o = (orc_t *)calloc(1, sizeof(orc_t));
  if (!o) return NULL;
  o->nx = nx; o->ny = ny; o->nz = nz;
  o->x0 = x0; o->y0 = y0; o->z0 = z0;
  o->gx = gx; o->gy = gy; o->gz = gz;
  /* src/logistics/initialize.F90:143-148: any absorbing axis makes all absorbing */
  if (bndx == 2 || bndy == 2 || bndz == 2) bndx = bndy = bndz = 2;
  o->bnd[0] = bndx; o->bnd[1] = bndy; o->bnd[2] = bndz;
  o->absorb_buff = absorb_buff;
  o->absorb_fixed = absorb_fixed;
  o->cc = (real)cc;
  o->ccinv = (real)1.0 / o->cc; /* initialize.F90:189 */
  o->sx = nx + 2 * NG; o->sy = ny + 2 * NG; o->sz = nz + 2 * NG;
  o->ncell = o->sx * o->sy * o->sz;
  for (int f = 0; f < F_COUNT; ++f) {
    o->f[f] = (real *)calloc((size_t)o->ncell, sizeof(real));
    if (!o->f[f]) return NULL;
  }
  return o;
}

void orc_destroy(orc_t *o) {
  if (!o) return;
  for (int f = 0; f < F_COUNT; ++f) free(o->f[f]);
  free(o);
}

real *orc_field(orc_t *o, int which) { return (which >= 0 && which < F_COUNT) ? o->f[which] : NULL; }
int64_t orc_ncell(const orc_t *o) { return o->ncell; }
int orc_real_size(void) { return (int)sizeof(real); }

void orc_set_bc(orc_t *o, int kind, const double *p, int n) {
  o->bc_kind = kind;
  memset(o->bcp, 0, sizeof(o->bcp));
  for (int i = 0; i < n && i < 16; ++i) o->bcp[i] = p[i];
}

/* --- src/algorithms/field_solver.F90:15-33  initRKstep ------------------ */
void orc_init_rk_step(orc_t *o) {
  const size_t nb = (size_t)o->ncell * sizeof(real);
  memcpy(o->f[F_ENX], o->f[F_EX], nb);
  memcpy(o->f[F_ENY], o->f[F_EY], nb);
  memcpy(o->f[F_ENZ], o->f[F_EZ], nb);
  memcpy(o->f[F_BNX], o->f[F_BX], nb);
  memcpy(o->f[F_BNY], o->f[F_BY], nb);
  memcpy(o->f[F_BNZ], o->f[F_BZ], nb);
  /* QNAN (src/defs.F90:8) takes the low word of the pair: it is 0.0 */
  for (int f = F_DEX; f <= F_RHO; ++f) memset(o->f[f], 0, nb);
}

/* --- field_solver.F90:44-63  computeRho --------------------------------- */
static void orc_compute_rho(orc_t *o) {
  real *ex = o->f[F_EX], *ey = o->f[F_EY], *ez = o->f[F_EZ], *rho = o->f[F_RHO];
#pragma omp parallel for schedule(static)
  for (int k = -1; k <= o->nz + 1; ++k)
    for (int j = -1; j <= o->ny + 1; ++j)
      for (int i = -1; i <= o->nx + 1; ++i)
        rho[IDX(o, i, j, k)] = (ex[IDX(o, i, j, k)] - ex[IDX(o, i - 1, j, k)]) +
                               (ey[IDX(o, i, j, k)] - ey[IDX(o, i, j - 1, k)]) +
                               (ez[IDX(o, i, j, k)] - ez[IDX(o, i, j, k - 1)]);
}

/* --- field_solver.F90:65-166  rk3Push ----------------------------------- */
static void orc_rk3_push(orc_t *o) {
  const real *ex = o->f[F_EX], *ey = o->f[F_EY], *ez = o->f[F_EZ];
  const real *bx = o->f[F_BX], *by = o->f[F_BY], *bz = o->f[F_BZ];
  const real *b0x = o->f[F_B0X], *b0y = o->f[F_B0Y], *b0z = o->f[F_B0Z];
  const real *rho = o->f[F_RHO];
  real *dex = o->f[F_DEX], *dey = o->f[F_DEY], *dez = o->f[F_DEZ];
  real *dbx = o->f[F_DBX], *dby = o->f[F_DBY], *dbz = o->f[F_DBZ];
  const real CC = o->cc;
#define A(arr, di, dj, dk) arr[IDX(o, i + (di), j + (dj), k + (dk))]
#pragma omp parallel for schedule(static)
  for (int k = 0; k < o->nz; ++k)
    for (int j = 0; j < o->ny; ++j)
      for (int i = 0; i < o->nx; ++i) {
        const int64_t c = IDX(o, i, j, k);
        real jx, jy, jz, corr;
        real intex, intey, intez, intbx, intby, intbz, bmag, emag, intrho;
        dbx[c] = CC * ((A(ey, 0, 0, 1) - A(ey, 0, 0, 0)) - (A(ez, 0, 1, 0) - A(ez, 0, 0, 0)));
        dby[c] = CC * ((A(ez, 1, 0, 0) - A(ez, 0, 0, 0)) - (A(ex, 0, 0, 1) - A(ex, 0, 0, 0)));
        dbz[c] = CC * ((A(ex, 0, 1, 0) - A(ex, 0, 0, 0)) - (A(ey, 1, 0, 0) - A(ey, 0, 0, 0)));
        dex[c] = -CC * (((A(by, 0, 0, 0) - A(by, 0, 0, -1)) - (A(bz, 0, 0, 0) - A(bz, 0, -1, 0))) -
                        ((A(b0y, 0, 0, 0) - A(b0y, 0, 0, -1)) - (A(b0z, 0, 0, 0) - A(b0z, 0, -1, 0))));
        dey[c] = -CC * (((A(bz, 0, 0, 0) - A(bz, -1, 0, 0)) - (A(bx, 0, 0, 0) - A(bx, 0, 0, -1))) -
                        ((A(b0z, 0, 0, 0) - A(b0z, -1, 0, 0)) - (A(b0x, 0, 0, 0) - A(b0x, 0, 0, -1))));
        dez[c] = -CC * (((A(bx, 0, 0, 0) - A(bx, 0, -1, 0)) - (A(by, 0, 0, 0) - A(by, -1, 0, 0))) -
                        ((A(b0x, 0, 0, 0) - A(b0x, 0, -1, 0)) - (A(b0y, 0, 0, 0) - A(b0y, -1, 0, 0))));
        /* X current (:90-111) */
        intrho = (real)0.5 * (A(rho, 1, 0, 0) + A(rho, 0, 0, 0));
        intex = A(ex, 0, 0, 0);
        intey = (real)0.25 * (A(ey, 0, 0, 0) + A(ey, 1, 0, 0) + A(ey, 0, -1, 0) + A(ey, 1, -1, 0));
        intez = (real)0.25 * (A(ez, 0, 0, 0) + A(ez, 1, 0, 0) + A(ez, 0, 0, -1) + A(ez, 1, 0, -1));
        intbx = (real)0.125 * (A(bx, 0, 0, 0) + A(bx, 0, -1, 0) + A(bx, 1, -1, 0) + A(bx, 1, 0, 0) +
                               A(bx, 0, 0, -1) + A(bx, 0, -1, -1) + A(bx, 1, -1, -1) + A(bx, 1, 0, -1));
        intby = (real)0.5 * (A(by, 0, 0, 0) + A(by, 0, 0, -1));
        intbz = (real)0.5 * (A(bz, 0, 0, 0) + A(bz, 0, -1, 0));
        bmag = intbx * intbx + intby * intby + intbz * intbz;
        emag = intex * intex + intey * intey + intez * intez;
        if (emag > bmag) {
          corr = (real)sqrt(bmag / emag);
          intey = intey * corr;
          intez = intez * corr;
        }
        jx = CC / (bmag + ORC_TINY) * (intrho * (intey * intbz - intby * intez));
        /* Y current (:113-134) */
        intrho = (real)0.5 * (A(rho, 0, 1, 0) + A(rho, 0, 0, 0));
        intex = (real)0.25 * (A(ex, 0, 0, 0) + A(ex, -1, 0, 0) + A(ex, 0, 1, 0) + A(ex, -1, 1, 0));
        intey = A(ey, 0, 0, 0);
        intez = (real)0.25 * (A(ez, 0, 0, 0) + A(ez, 0, 1, 0) + A(ez, 0, 0, -1) + A(ez, 0, 1, -1));
        intbx = (real)0.5 * (A(bx, 0, 0, 0) + A(bx, 0, 0, -1));
        intby = (real)0.125 * (A(by, 0, 0, 0) + A(by, -1, 0, 0) + A(by, -1, 1, 0) + A(by, 0, 1, 0) +
                               A(by, 0, 0, -1) + A(by, -1, 0, -1) + A(by, -1, 1, -1) + A(by, 0, 1, -1));
        intbz = (real)0.5 * (A(bz, 0, 0, 0) + A(bz, -1, 0, 0));
        bmag = intbx * intbx + intby * intby + intbz * intbz;
        emag = intex * intex + intey * intey + intez * intez;
        if (emag > bmag) {
          corr = (real)sqrt(bmag / emag);
          intex = intex * corr;
          intez = intez * corr;
        }
        jy = CC / (bmag + ORC_TINY) * (intrho * (intez * intbx - intex * intbz));
        /* Z current (:136-157) */
        intrho = (real)0.5 * (A(rho, 0, 0, 0) + A(rho, 0, 0, 1));
        intex = (real)0.25 * (A(ex, 0, 0, 0) + A(ex, -1, 0, 0) + A(ex, 0, 0, 1) + A(ex, -1, 0, 1));
        intey = (real)0.25 * (A(ey, 0, 0, 0) + A(ey, 0, -1, 0) + A(ey, 0, 0, 1) + A(ey, 0, -1, 1));
        intez = A(ez, 0, 0, 0);
        intbx = (real)0.5 * (A(bx, 0, 0, 0) + A(bx, 0, -1, 0));
        intby = (real)0.5 * (A(by, 0, 0, 0) + A(by, -1, 0, 0));
        intbz = (real)0.125 * (A(bz, 0, 0, 0) + A(bz, -1, 0, 0) + A(bz, -1, -1, 0) + A(bz, 0, -1, 0) +
                               A(bz, 0, 0, 1) + A(bz, -1, 0, 1) + A(bz, -1, -1, 1) + A(bz, 0, -1, 1));
        bmag = intbx * intbx + intby * intby + intbz * intbz;
        emag = intex * intex + intey * intey + intez * intez;
        if (emag > bmag) {
          corr = (real)sqrt(bmag / emag);
          intex = intex * corr;
          intey = intey * corr;
        }
        jz = CC / (bmag + ORC_TINY) * (intrho * (intex * intby - intbx * intey));
        /* :160-162 */
        dex[c] = dex[c] - jx;
        dey[c] = dey[c] - jy;
        dez[c] = dez[c] - jz;
      }
#undef A
}

/* --- field_solver.F90:168-271  rk3Update -------------------------------- */
static inline real cube(real d) { return d * d * d; }

static void orc_rk3_update(orc_t *o, real c1, real c2, real c3) {
  real *ex = o->f[F_EX], *ey = o->f[F_EY], *ez = o->f[F_EZ];
  real *bx = o->f[F_BX], *by = o->f[F_BY], *bz = o->f[F_BZ];
  const real *enx = o->f[F_ENX], *eny = o->f[F_ENY], *enz = o->f[F_ENZ];
  const real *bnx = o->f[F_BNX], *bny = o->f[F_BNY], *bnz = o->f[F_BNZ];
  const real *dex = o->f[F_DEX], *dey = o->f[F_DEY], *dez = o->f[F_DEZ];
  const real *dbx = o->f[F_DBX], *dby = o->f[F_DBY], *dbz = o->f[F_DBZ];
  const int absorbing = (o->bnd[0] == 2);
  const real xc = (real)0.5 * (real)o->gx, yc = (real)0.5 * (real)o->gy, zc = (real)0.5 * (real)o->gz;
  real rmax = xc < yc ? xc : yc;
  rmax = rmax < zc ? rmax : zc;
  const real rabsorb = rmax - (real)o->absorb_buff;
  const real buff = (real)o->absorb_buff;
  const int fixed = o->absorb_fixed;
#pragma omp parallel for schedule(static)
  for (int k = 0; k < o->nz; ++k)
    for (int j = 0; j < o->ny; ++j)
      for (int i = 0; i < o->nx; ++i) {
        const int64_t c = IDX(o, i, j, k);
        ex[c] = c1 * enx[c] + c2 * ex[c] + c3 * dex[c];
        ey[c] = c1 * eny[c] + c2 * ey[c] + c3 * dey[c];
        ez[c] = c1 * enz[c] + c2 * ez[c] + c3 * dez[c];
        bx[c] = c1 * bnx[c] + c2 * bx[c] + c3 * dbx[c];
        by[c] = c1 * bny[c] + c2 * by[c] + c3 * dby[c];
        bz[c] = c1 * bnz[c] + c2 * bz[c] + c3 * dbz[c];
        if (!absorbing) continue;
        /* :210-212 -- as coded, jg is built from i (quirk); fixed mode uses j */
        const real ig = (real)(i + o->x0);
        const real jg = fixed ? (real)(j + o->y0) : (real)(i + o->y0);
        const real kg = (real)(k + o->z0);
        real r = (real)sqrt((ig - xc) * (ig - xc) + (jg - yc) * (jg - yc) + (kg - zc) * (kg - zc));
        if (r > rabsorb - (real)1.0) {
          real d;
#define DAMP(arr, ox, oy, oz)                                                                   \
  r = (real)sqrt((ig + (ox) - xc) * (ig + (ox) - xc) + (jg + (oy) - yc) * (jg + (oy) - yc) + \
                 (kg + (oz) - zc) * (kg + (oz) - zc));                                        \
  d = (r - rabsorb) / buff;                                                                   \
  if (fixed && d > (real)1.0) d = (real)1.0;                                                  \
  if (d > (real)0.0) arr[c] = ((real)1.0 - cube(d)) * arr[c];
          const real h = (real)0.5, z = (real)0.0;
          DAMP(ex, h, z, z)
          DAMP(ey, z, h, z)
          DAMP(ez, z, z, h)
          DAMP(bx, z, h, h)
          DAMP(by, h, z, h)
          DAMP(bz, h, h, z)
#undef DAMP
        }
      }
}

/* --- field_solver.F90:35-42  rk3Step ------------------------------------ */
void orc_rk3_step(orc_t *o, double c1, double c2, double c3) {
  orc_compute_rho(o);
  orc_rk3_push(o);
  orc_rk3_update(o, (real)c1, (real)c2, (real)c3);
}

/* --- field_solver.F90:273-335  cleanEpar --------------------------------
 * As coded the projection divides by |B|^2 with no floor, so an exact null
 * gives NaN.  In fixed mode (absorb_fixed) the absorber zeroes the fields
 * beyond r_max and so creates exact nulls; there the projection is skipped
 * where |B|^2 == 0 and E is left as it is. */
void orc_clean_epar(orc_t *o) {
  const int guard = o->absorb_fixed;
  real *ex = o->f[F_EX], *ey = o->f[F_EY], *ez = o->f[F_EZ];
  const real *bx = o->f[F_BX], *by = o->f[F_BY], *bz = o->f[F_BZ];
  real *dex = o->f[F_DEX], *dey = o->f[F_DEY], *dez = o->f[F_DEZ];
#define A(arr, di, dj, dk) arr[IDX(o, i + (di), j + (dj), k + (dk))]
#pragma omp parallel for schedule(static)
  for (int k = 0; k < o->nz; ++k)
    for (int j = 0; j < o->ny; ++j)
      for (int i = 0; i < o->nx; ++i) {
        const int64_t c = IDX(o, i, j, k);
        real intex, intey, intez, intbx, intby, intbz, bmag, epar;
        intex = A(ex, 0, 0, 0);
        intey = (real)0.25 * (A(ey, 0, 0, 0) + A(ey, 1, 0, 0) + A(ey, 0, -1, 0) + A(ey, 1, -1, 0));
        intez = (real)0.25 * (A(ez, 0, 0, 0) + A(ez, 1, 0, 0) + A(ez, 0, 0, -1) + A(ez, 1, 0, -1));
        intbx = (real)0.125 * (A(bx, 0, 0, 0) + A(bx, 0, -1, 0) + A(bx, 1, -1, 0) + A(bx, 1, 0, 0) +
                               A(bx, 0, 0, -1) + A(bx, 0, -1, -1) + A(bx, 1, -1, -1) + A(bx, 1, 0, -1));
        intby = (real)0.5 * (A(by, 0, 0, 0) + A(by, 0, 0, -1));
        intbz = (real)0.5 * (A(bz, 0, 0, 0) + A(bz, 0, -1, 0));
        bmag = intbx * intbx + intby * intby + intbz * intbz;
        epar = (intex * intbx + intey * intby + intez * intbz);
        dex[c] = (guard && bmag == (real)0) ? A(ex, 0, 0, 0) : A(ex, 0, 0, 0) - epar * intbx / bmag;

        intex = (real)0.25 * (A(ex, 0, 0, 0) + A(ex, -1, 0, 0) + A(ex, 0, 1, 0) + A(ex, -1, 1, 0));
        intey = A(ey, 0, 0, 0);
        intez = (real)0.25 * (A(ez, 0, 0, 0) + A(ez, 0, 1, 0) + A(ez, 0, 0, -1) + A(ez, 0, 1, -1));
        intbx = (real)0.5 * (A(bx, 0, 0, 0) + A(bx, 0, 0, -1));
        intby = (real)0.125 * (A(by, 0, 0, 0) + A(by, -1, 0, 0) + A(by, -1, 1, 0) + A(by, 0, 1, 0) +
                               A(by, 0, 0, -1) + A(by, -1, 0, -1) + A(by, -1, 1, -1) + A(by, 0, 1, -1));
        intbz = (real)0.5 * (A(bz, 0, 0, 0) + A(bz, -1, 0, 0));
        bmag = intbx * intbx + intby * intby + intbz * intbz;
        epar = (intex * intbx + intey * intby + intez * intbz);
        dey[c] = (guard && bmag == (real)0) ? A(ey, 0, 0, 0) : A(ey, 0, 0, 0) - epar * intby / bmag;

        intex = (real)0.25 * (A(ex, 0, 0, 0) + A(ex, -1, 0, 0) + A(ex, 0, 0, 1) + A(ex, -1, 0, 1));
        intey = (real)0.25 * (A(ey, 0, 0, 0) + A(ey, 0, -1, 0) + A(ey, 0, 0, 1) + A(ey, 0, -1, 1));
        intez = A(ez, 0, 0, 0);
        intbx = (real)0.5 * (A(bx, 0, 0, 0) + A(bx, 0, -1, 0));
        intby = (real)0.5 * (A(by, 0, 0, 0) + A(by, -1, 0, 0));
        intbz = (real)0.125 * (A(bz, 0, 0, 0) + A(bz, -1, 0, 0) + A(bz, -1, -1, 0) + A(bz, 0, -1, 0) +
                               A(bz, 0, 0, 1) + A(bz, -1, 0, 1) + A(bz, -1, -1, 1) + A(bz, 0, -1, 1));
        bmag = intbx * intbx + intby * intby + intbz * intbz;
        epar = (intex * intbx + intey * intby + intez * intbz);
        dez[c] = (guard && bmag == (real)0) ? A(ez, 0, 0, 0) : A(ez, 0, 0, 0) - epar * intbz / bmag;
      }
  /* :332-334 copy back over the interior */
#pragma omp parallel for schedule(static)
  for (int k = 0; k < o->nz; ++k)
    for (int j = 0; j < o->ny; ++j)
      for (int i = 0; i < o->nx; ++i) {
        const int64_t c = IDX(o, i, j, k);
        ex[c] = dex[c];
        ey[c] = dey[c];
        ez[c] = dez[c];
      }
#undef A
}

/* --- src/logistics/exchange_fields.F90:11-197 with one rank -------------
 * With a single meshblock, neighbour(d1,d2,d3) exists iff every axis with
 * d != 0 is periodic (src/tools/helpers.F90:22-46), and the neighbour is the
 * block itself; the message received from direction d is the slab the block
 * sends towards -d.  So the ghost slab in direction d receives the interior
 * slab on the opposite side: ghost index -g <- s-g and s-1+g <- g-1. */
static int orc_range(int d, int s, int recv, int *lo, int *hi) {
  if (d == 0) { *lo = 0; *hi = s - 1; }
  else if (d == -1) { if (recv) { *lo = -NG; *hi = -1; } else { *lo = 0; *hi = NG - 1; } }
  else { if (recv) { *lo = s; *hi = s + NG - 1; } else { *lo = s - NG; *hi = s - 1; } }
  return 0;
}

void orc_exchange_fields(orc_t *o, int exchange_b0) {
  const int nf = exchange_b0 ? 9 : 6;
  static const int fl[9] = {F_EX, F_EY, F_EZ, F_BX, F_BY, F_BZ, F_B0X, F_B0Y, F_B0Z};
  for (int d1 = -1; d1 <= 1; ++d1)
    for (int d2 = -1; d2 <= 1; ++d2)
      for (int d3 = -1; d3 <= 1; ++d3) {
        if (d1 == 0 && d2 == 0 && d3 == 0) continue;
        if ((d1 && o->bnd[0] != 1) || (d2 && o->bnd[1] != 1) || (d3 && o->bnd[2] != 1)) continue;
        int ri0, ri1, rj0, rj1, rk0, rk1, si0, si1, sj0, sj1, sk0, sk1;
        orc_range(d1, o->nx, 1, &ri0, &ri1);
        orc_range(d2, o->ny, 1, &rj0, &rj1);
        orc_range(d3, o->nz, 1, &rk0, &rk1);
        /* the sender (this same block, seen as neighbour d) packs towards -d */
        orc_range(-d1, o->nx, 0, &si0, &si1);
        orc_range(-d2, o->ny, 0, &sj0, &sj1);
        orc_range(-d3, o->nz, 0, &sk0, &sk1);
        for (int f = 0; f < nf; ++f) {
          real *a = o->f[fl[f]];
          for (int k = rk0; k <= rk1; ++k)
            for (int j = rj0; j <= rj1; ++j)
              for (int i = ri0; i <= ri1; ++i)
                a[IDX(o, i, j, k)] = a[IDX(o, si0 + (i - ri0), sj0 + (j - rj0), sk0 + (k - rk0))];
        }
      }
}

/* --- user/user_loop.F90:101-164  userFieldBoundaryConditions ------------
 * bcp: 0 xc, 1 yc1, 2 yc2, 3 zc, 4 bnorm, 5 r0, 6 beta_twist, 7 t0,
 *      8 twist (0: as shipped, floor/top copy; 1: the commented-out
 *        footpoint-rotation driver of user_loop.F90:122-150 on the floor
 *        ghosts, with the tanh(time/t0) ramp switched on; 2: the driven loop
 *        described in the function). */
static void orc_bc_loop(orc_t *o, int step) {
  real *ex = o->f[F_EX], *ey = o->f[F_EY], *bz = o->f[F_BZ];
  const real xc = (real)o->bcp[0], yc1 = (real)o->bcp[1], yc2 = (real)o->bcp[2];
  const real zc = (real)o->bcp[3], bnorm = (real)o->bcp[4], r0 = (real)o->bcp[5];
  const real beta = (real)o->bcp[6], t0 = (real)o->bcp[7];
  /* twist 0: floor/top copy as shipped; 1: the commented footpoint driver on
   * the floor ghosts (z < 0); 2: builder-defined driven loop in the form of
   * user_uniform.F90:113-142 -- the driver's ex, ey on z <= 0 (the floor
   * plane included) and a conducting top (ex = ey = 0, bz = the potential
   * field).  Unlike user_uniform.F90:134-137, bz of the floor plane itself
   * is not pinned: it follows Faraday's law under the driver's E, so the
   * floor cells keep div_h B (pinning it while the rotation's E has a
   * non-zero discrete curl changes div_h B there every stage); bz is
   * pinned on the ghost planes z < 0 only. */
  const int twist = (int)o->bcp[8];
  const real time = (real)step;
  const real CC = o->cc;
  for (int i = -NG; i <= o->nx - 1 + NG; ++i) {
    const real xg = (real)(i + o->x0);
    for (int j = -NG; j <= o->ny - 1 + NG; ++j) {
      const real yg = (real)(j + o->y0);
      for (int k = -NG; k <= o->nz - 1 + NG; ++k) {
        const real zg = (real)(k + o->z0);
        const int64_t c = IDX(o, i, j, k);
        if (zg < (real)0 || (twist == 2 && zg <= (real)0)) {
          if (twist) {
            real rx, ry1, ry2, rz, rr1, rr2, bzl, omega, ee0;
            rx = xg + (real)0.5 - xc; ry1 = yg - yc1; ry2 = yg - yc2; rz = zg - zc;
            rr1 = (real)sqrt(rx * rx + ry1 * ry1 + rz * rz);
            rr2 = (real)sqrt(rx * rx + ry2 * ry2 + rz * rz);
            bzl = bnorm * (rz / cube(rr1) - rz / cube(rr2));
            omega = (real)exp(-rr1 / r0) * beta * CC / r0;
            ee0 = -omega * bzl * (real)tanh(time / t0);
            ex[c] = rx * ee0;
            rx = xg - xc; ry1 = yg + (real)0.5 - yc1; ry2 = yg + (real)0.5 - yc2; rz = zg - zc;
            rr1 = (real)sqrt(rx * rx + ry1 * ry1 + rz * rz);
            rr2 = (real)sqrt(rx * rx + ry2 * ry2 + rz * rz);
            bzl = bnorm * (rz / cube(rr1) - rz / cube(rr2));
            omega = (real)exp(-rr1 / r0) * beta * CC / r0;
            ee0 = -omega * bzl * (real)tanh(time / t0);
            ey[c] = ry1 * ee0;
            rx = xg + (real)0.5 - xc; ry1 = yg + (real)0.5 - yc1; ry2 = yg + (real)0.5 - yc2; rz = zg - zc;
            rr1 = (real)sqrt(rx * rx + ry1 * ry1 + rz * rz);
            rr2 = (real)sqrt(rx * rx + ry2 * ry2 + rz * rz);
            if (zg < (real)0) bz[c] = bnorm * (rz / cube(rr1) - rz / cube(rr2));
          } else {
            ex[c] = ex[IDX(o, i, j, 0)];
            ey[c] = ey[IDX(o, i, j, 0)];
            bz[c] = bz[IDX(o, i, j, 0)];
          }
        }
        if (zg >= (real)o->gz) {
          if (twist == 2) {
            const real rx = xg + (real)0.5 - xc, ry1 = yg + (real)0.5 - yc1;
            const real ry2 = yg + (real)0.5 - yc2, rz = zg - zc;
            const real rr1 = (real)sqrt(rx * rx + ry1 * ry1 + rz * rz);
            const real rr2 = (real)sqrt(rx * rx + ry2 * ry2 + rz * rz);
            bz[c] = bnorm * (rz / cube(rr1) - rz / cube(rr2));
            ex[c] = (real)0;
            ey[c] = (real)0;
          } else {
            ex[c] = ex[IDX(o, i, j, o->nz - 1)];
            ey[c] = ey[IDX(o, i, j, o->nz - 1)];
            bz[c] = bz[IDX(o, i, j, o->nz - 1)];
          }
        }
      }
    }
  }
}

/* --- user/user_uniform.F90:77-147  userFieldBoundaryConditions ----------
 * bcp: 0 xc, 1 yc, 2 r0 (already multiplied by L0), 3 beta_twist, 4 t0, 5 bnorm */
static void orc_bc_uniform(orc_t *o, int step) {
  real *ex = o->f[F_EX], *ey = o->f[F_EY], *bz = o->f[F_BZ];
  const real xc = (real)o->bcp[0], yc = (real)o->bcp[1], r0 = (real)o->bcp[2];
  const real beta = (real)o->bcp[3], t0 = (real)o->bcp[4], bnorm = (real)o->bcp[5];
  const real time = (real)step, CC = o->cc;
  for (int i = -NG; i <= o->nx - 1 + NG; ++i) {
    const real xg = (real)(i + o->x0);
    for (int j = -NG; j <= o->ny - 1 + NG; ++j) {
      const real yg = (real)(j + o->y0);
      for (int k = -NG; k <= o->nz - 1 + NG; ++k) {
        const real zg = (real)(k + o->z0);
        const int64_t c = IDX(o, i, j, k);
        real rx, ry, rr, omega, ee0;
        if (zg <= (real)0) {
          rx = xg + (real)0.5 - xc; ry = yg - yc;
          rr = (real)sqrt(rx * rx + ry * ry + ORC_TINY);
          omega = (real)exp(-rr / r0) * beta * CC / r0;
          ee0 = -omega * bnorm * (real)tanh(time / t0);
          ex[c] = rx * ee0;
          rx = xg - xc; ry = yg + (real)0.5 - yc;
          rr = (real)sqrt(rx * rx + ry * ry + ORC_TINY);
          omega = (real)exp(-rr / r0) * beta * CC / r0;
          ee0 = -omega * bnorm * (real)tanh(time / t0);
          ey[c] = ry * ee0;
          rx = xg + (real)0.5 - xc; ry = yg + (real)0.5 - yc;
          rr = (real)sqrt(rx * rx + ry * ry + ORC_TINY);
          bz[c] = bnorm / ((real)1.0 + (rr / r0));
        }
        if (zg >= (real)o->gz) {
          rx = xg + (real)0.5 - xc; ry = yg + (real)0.5 - yc;
          rr = (real)sqrt(rx * rx + ry * ry + ORC_TINY);
          bz[c] = bnorm / ((real)1.0 + (rr / r0));
          ex[c] = (real)0.0;
          ey[c] = (real)0.0;
        }
      }
    }
  }
}

/* --- user/user_psr.F90:345-383  getDipole / shape ------------------------
 * bcp: 0 psr_radius, 1 psr_angle (rad), 2 psr_omega, 3 xc_g, 4 yc_g, 5 zc_g */
static void orc_dipole(const orc_t *o, int step, real xg, real yg, real zg, real *obx, real *oby,
                       real *obz) {
  const real R = (real)o->bcp[0], ang = (real)o->bcp[1], om = (real)o->bcp[2];
  const real xc = (real)o->bcp[3], yc = (real)o->bcp[4], zc = (real)o->bcp[5];
  real phase = om * (real)step;
  real nx = xg - xc, ny = yg - yc, nz = zg - zc;
  real rr = (real)1.0 / (real)sqrt(nx * nx + ny * ny + nz * nz);
  nx = nx * rr; ny = ny * rr; nz = nz * rr;
  rr = cube(rr);
  const real offset = (real)0.0;
  real mux = cube(R) * (real)sin(ang) * (real)cos(phase + offset);
  real muy = cube(R) * (real)sin(ang) * (real)sin(phase + offset);
  real muz = cube(R) * (real)cos(ang);
  real mdn = mux * nx + muy * ny + muz * nz;
  *obx = ((real)3.0 * nx * mdn - mux) * rr;
  *oby = ((real)3.0 * ny * mdn - muy) * rr;
  *obz = ((real)3.0 * nz * mdn - muz) * rr;
}

static inline real orc_shape(real rad, real rad0) {
  const real del = (real)1.0;
  return (real)0.5 * ((real)1.0 - (real)tanh((rad - rad0) / del));
}

/* --- user/user_psr.F90:72-332  userFieldBoundaryConditions -------------- */
static void orc_bc_psr(orc_t *o, int step) {
  real *ex = o->f[F_EX], *ey = o->f[F_EY], *ez = o->f[F_EZ];
  real *bx = o->f[F_BX], *by = o->f[F_BY], *bz = o->f[F_BZ];
  const real R = (real)o->bcp[0], om = (real)o->bcp[2];
  const real xc = (real)o->bcp[3], yc = (real)o->bcp[4], zc = (real)o->bcp[5];
  const real CCINV = o->ccinv;
  const real ssq = (R * (real)2) * (R * (real)2);
  /* :93-96 does the block touch the supersphere */
  {
    real x_ = (real)(o->x0 - NG), xh = (real)(o->x0 + o->nx - 1 + NG);
    real y_ = (real)(o->y0 - NG), yh = (real)(o->y0 + o->ny - 1 + NG);
    real z_ = (real)(o->z0 - NG), zh = (real)(o->z0 + o->nz - 1 + NG);
    x_ = x_ > (xh < xc ? xh : xc) ? x_ : (xh < xc ? xh : xc);
    y_ = y_ > (yh < yc ? yh : yc) ? y_ : (yh < yc ? yh : yc);
    z_ = z_ > (zh < zc ? zh : zc) ? z_ : (zh < zc ? zh : zc);
    if (!((x_ - xc) * (x_ - xc) + (y_ - yc) * (y_ - yc) + (z_ - zc) * (z_ - zc) < ssq)) return;
  }
  const real shift_B = (real)4.0, shift_E = (real)0.0;
  const real sEpar = (real)0.5, sEperp = (real)0.25, sBperp = (real)0.5, sBpar = (real)0.25;
  real *nw[6];
  for (int f = 0; f < 6; ++f) nw[f] = (real *)calloc((size_t)o->ncell, sizeof(real));
#define A(arr, di, dj, dk) arr[IDX(o, i + (di), j + (dj), k + (dk))]
#pragma omp parallel for schedule(static)
  for (int k = 0; k < o->nz; ++k)
    for (int j = 0; j < o->ny; ++j)
      for (int i = 0; i < o->nx; ++i) {
        const int ig = i + o->x0, jg = j + o->y0, kg = k + o->z0;
        if (!(((real)ig - xc) * ((real)ig - xc) + ((real)jg - yc) * ((real)jg - yc) +
                  ((real)kg - zc) * ((real)kg - zc) < ssq))
          continue;
        const int64_t c = IDX(o, i, j, k);
        real rx, ry, rz, rr_sqr, s, scale, bdx, bdy, bdz, bidr, bddr, v;
        /* B_x :157-174 */
        rx = (real)ig - xc; ry = (real)jg + (real)0.5 - yc; rz = (real)kg + (real)0.5 - zc;
        rr_sqr = rx * rx + ry * ry + rz * rz;
        bidr = A(bx, 0, 0, 0) * rx +
               (real)0.25 * (A(by, 0, 0, 0) + A(by, 0, 1, 0) + A(by, -1, 0, 0) + A(by, -1, 1, 0)) * ry +
               (real)0.25 * (A(bz, 0, 0, 0) + A(bz, 0, 0, 1) + A(bz, -1, 0, 0) + A(bz, -1, 0, 1)) * rz;
        orc_dipole(o, step, rx + xc, ry + yc, rz + zc, &bdx, &bdy, &bdz);
        bddr = bdx * rx + bdy * ry + bdz * rz;
        scale = (real)1.0 / sBperp;
        s = orc_shape((real)sqrt(rr_sqr) * scale, (R - shift_B) * scale);
        v = (bidr - bddr) * (rx / rr_sqr) * ((real)1.0 - s);
        scale = (real)1.0 / sBpar;
        s = orc_shape((real)sqrt(rr_sqr) * scale, (R - shift_B) * scale);
        nw[3][c] = v + bdx + ((A(bx, 0, 0, 0) - bidr * rx / rr_sqr) - (bdx - bddr * rx / rr_sqr)) * ((real)1.0 - s);
        /* B_y :177-194 */
        rx = (real)ig + (real)0.5 - xc; ry = (real)jg - yc; rz = (real)kg + (real)0.5 - zc;
        rr_sqr = rx * rx + ry * ry + rz * rz;
        bidr = (real)0.25 * (A(bx, 0, 0, 0) + A(bx, 1, 0, 0) + A(bx, 0, -1, 0) + A(bx, 1, -1, 0)) * rx +
               A(by, 0, 0, 0) * ry +
               (real)0.25 * (A(bz, 0, 0, 0) + A(bz, 0, 0, 1) + A(bz, 0, -1, 0) + A(bz, 0, -1, 1)) * rz;
        orc_dipole(o, step, rx + xc, ry + yc, rz + zc, &bdx, &bdy, &bdz);
        bddr = bdx * rx + bdy * ry + bdz * rz;
        scale = (real)1.0 / sBperp;
        s = orc_shape((real)sqrt(rr_sqr) * scale, (R - shift_B) * scale);
        v = (bidr - bddr) * (ry / rr_sqr) * ((real)1.0 - s);
        scale = (real)1.0 / sBpar;
        s = orc_shape((real)sqrt(rr_sqr) * scale, (R - shift_B) * scale);
        nw[4][c] = v + bdy + ((A(by, 0, 0, 0) - bidr * ry / rr_sqr) - (bdy - bddr * ry / rr_sqr)) * ((real)1.0 - s);
        /* B_z :197-214 */
        rx = (real)ig + (real)0.5 - xc; ry = (real)jg + (real)0.5 - yc; rz = (real)kg - zc;
        rr_sqr = rx * rx + ry * ry + rz * rz;
        bidr = (real)0.25 * (A(bx, 0, 0, 0) + A(bx, 1, 0, 0) + A(bx, 0, 0, -1) + A(bx, 1, 0, -1)) * rx +
               (real)0.25 * (A(by, 0, 0, 0) + A(by, 0, 1, 0) + A(by, 0, 0, -1) + A(by, 0, 1, -1)) * ry +
               A(bz, 0, 0, 0) * rz;
        orc_dipole(o, step, rx + xc, ry + yc, rz + zc, &bdx, &bdy, &bdz);
        bddr = bdx * rx + bdy * ry + bdz * rz;
        scale = (real)1.0 / sBperp;
        s = orc_shape((real)sqrt(rr_sqr) * scale, (R - shift_B) * scale);
        v = (bidr - bddr) * (rz / rr_sqr) * ((real)1.0 - s);
        scale = (real)1.0 / sBpar;
        s = orc_shape((real)sqrt(rr_sqr) * scale, (R - shift_B) * scale);
        nw[5][c] = v + bdz + ((A(bz, 0, 0, 0) - bidr * rz / rr_sqr) - (bdz - bddr * rz / rr_sqr)) * ((real)1.0 - s);

        real vx, vy, vz, edx, edy, edz, eidr, eddr;
        /* E_x :219-243 */
        rx = (real)ig + (real)0.5 - xc; ry = (real)jg - yc; rz = (real)kg - zc;
        rr_sqr = rx * rx + ry * ry + rz * rz;
        eidr = A(ex, 0, 0, 0) * rx +
               (real)0.25 * (A(ey, 0, 0, 0) + A(ey, 1, 0, 0) + A(ey, 0, -1, 0) + A(ey, 1, -1, 0)) * ry +
               (real)0.25 * (A(ez, 0, 0, 0) + A(ez, 1, 0, 0) + A(ez, 0, 0, -1) + A(ez, 1, 0, -1)) * rz;
        orc_dipole(o, step, rx + xc, ry + yc, rz + zc, &bdx, &bdy, &bdz);
        vx = -om * ry; vy = om * rx; vz = (real)0.0;
        edx = -(vy * bdz - vz * bdy) * CCINV;
        edy = (vx * bdz - vz * bdx) * CCINV;
        edz = -(vx * bdy - vy * bdx) * CCINV;
        eddr = edx * rx + edy * ry + edz * rz;
        scale = (real)1.0 / sEpar;
        s = orc_shape((real)sqrt(rr_sqr) * scale, (R - shift_E) * scale);
        v = edx + ((A(ex, 0, 0, 0) - eidr * rx / rr_sqr) - (edx - eddr * rx / rr_sqr)) * ((real)1.0 - s);
        scale = (real)1.0 / sEperp;
        s = orc_shape((real)sqrt(rr_sqr) * scale, (R - shift_E) * scale);
        nw[0][c] = v + (eidr - eddr) * (rx / rr_sqr) * ((real)1.0 - s);
        /* E_y :246-270 */
        rx = (real)ig - xc; ry = (real)jg + (real)0.5 - yc; rz = (real)kg - zc;
        rr_sqr = rx * rx + ry * ry + rz * rz;
        eidr = (real)0.25 * (A(ex, 0, 0, 0) + A(ex, 0, 1, 0) + A(ex, -1, 0, 0) + A(ex, -1, 1, 0)) * rx +
               A(ey, 0, 0, 0) * ry +
               (real)0.25 * (A(ez, 0, 0, 0) + A(ez, 0, 1, 0) + A(ez, 0, 0, -1) + A(ez, 0, 1, -1)) * rz;
        orc_dipole(o, step, rx + xc, ry + yc, rz + zc, &bdx, &bdy, &bdz);
        vx = -om * ry; vy = om * rx; vz = (real)0.0;
        edx = -(vy * bdz - vz * bdy) * CCINV;
        edy = (vx * bdz - vz * bdx) * CCINV;
        edz = -(vx * bdy - vy * bdx) * CCINV;
        eddr = edx * rx + edy * ry + edz * rz;
        scale = (real)1.0 / sEpar;
        s = orc_shape((real)sqrt(rr_sqr) * scale, (R - shift_E) * scale);
        v = edy + ((A(ey, 0, 0, 0) - eidr * ry / rr_sqr) - (edy - eddr * ry / rr_sqr)) * ((real)1.0 - s);
        scale = (real)1.0 / sEperp;
        s = orc_shape((real)sqrt(rr_sqr) * scale, (R - shift_E) * scale);
        nw[1][c] = v + (eidr - eddr) * (ry / rr_sqr) * ((real)1.0 - s);
        /* E_z :273-297 */
        rx = (real)ig - xc; ry = (real)jg - yc; rz = (real)kg + (real)0.5 - zc;
        rr_sqr = rx * rx + ry * ry + rz * rz;
        eidr = (real)0.25 * (A(ex, 0, 0, 0) + A(ex, 0, 0, 1) + A(ex, -1, 0, 0) + A(ex, -1, 0, 1)) * rx +
               (real)0.25 * (A(ey, 0, 0, 0) + A(ey, 0, -1, 0) + A(ey, 0, 0, 1) + A(ey, 0, -1, 1)) * ry +
               A(ez, 0, 0, 0) * rz;
        orc_dipole(o, step, rx + xc, ry + yc, rz + zc, &bdx, &bdy, &bdz);
        vx = -om * ry; vy = om * rx; vz = (real)0.0;
        edx = -(vy * bdz - vz * bdy) * CCINV;
        edy = (vx * bdz - vz * bdx) * CCINV;
        edz = -(vx * bdy - vy * bdx) * CCINV;
        eddr = edx * rx + edy * ry + edz * rz;
        scale = (real)1.0 / sEpar;
        s = orc_shape((real)sqrt(rr_sqr) * scale, (R - shift_E) * scale);
        v = edz + ((A(ez, 0, 0, 0) - eidr * rz / rr_sqr) - (edz - eddr * rz / rr_sqr)) * ((real)1.0 - s);
        scale = (real)1.0 / sEperp;
        s = orc_shape((real)sqrt(rr_sqr) * scale, (R - shift_E) * scale);
        nw[2][c] = v + (eidr - eddr) * (rz / rr_sqr) * ((real)1.0 - s);
      }
#undef A
  /* :303-325 copy back */
  real *dst[6] = {ex, ey, ez, bx, by, bz};
#pragma omp parallel for schedule(static)
  for (int k = 0; k < o->nz; ++k)
    for (int j = 0; j < o->ny; ++j)
      for (int i = 0; i < o->nx; ++i) {
        const int ig = i + o->x0, jg = j + o->y0, kg = k + o->z0;
        if (!(((real)ig - xc) * ((real)ig - xc) + ((real)jg - yc) * ((real)jg - yc) +
                  ((real)kg - zc) * ((real)kg - zc) < ssq))
          continue;
        const int64_t c = IDX(o, i, j, k);
        for (int f = 0; f < 6; ++f) dst[f][c] = nw[f][c];
      }
  for (int f = 0; f < 6; ++f) free(nw[f]);
}

/* userFieldBoundaryConditions(step) for the registered problem */
void orc_user_bc(orc_t *o, int step) {
  switch (o->bc_kind) {
    case BC_LOOP: orc_bc_loop(o, step); break;
    case BC_PSR: orc_bc_psr(o, step); break;
    case BC_UNIFORM: orc_bc_uniform(o, step); break;
    default: break;
  }
}

/* --- src/tristanmainloop.F90:40-42  pre-loop ghost fill + BC(0) --------- */
void orc_preloop(orc_t *o) {
  orc_exchange_fields(o, 1);
  orc_user_bc(o, 0);
}

/* --- src/tristanmainloop.F90:50-83  one timestep ------------------------
 * coef selects the stage coefficients: the caller passes them already in
 * the working precision (3/4, 1/4, 1/3, 2/3). */
void orc_step(orc_t *o, int timestep, const double *coef) {
  orc_init_rk_step(o);
  orc_rk3_step(o, coef[0], coef[1], coef[2]);
  orc_user_bc(o, timestep);
  orc_exchange_fields(o, 0);
  orc_clean_epar(o);
  orc_exchange_fields(o, 0);
  orc_rk3_step(o, coef[3], coef[4], coef[5]);
  orc_user_bc(o, timestep);
  orc_exchange_fields(o, 0);
  orc_clean_epar(o);
  orc_exchange_fields(o, 0);
  orc_rk3_step(o, coef[6], coef[7], coef[8]);
  orc_user_bc(o, timestep);
  orc_exchange_fields(o, 0);
  orc_clean_epar(o);
  orc_exchange_fields(o, 0);
}

void orc_run(orc_t *o, int first_step, int nsteps, const double *coef, double *energy_hist) {
  for (int n = 0; n < nsteps; ++n) {
    orc_step(o, first_step + n, coef);
    if (energy_hist) {
      double e2 = 0.0, b2 = 0.0;
      for (int k = 0; k < o->nz; ++k)
        for (int j = 0; j < o->ny; ++j)
          for (int i = 0; i < o->nx; ++i) {
            const int64_t c = IDX(o, i, j, k);
            const double a = o->f[F_EX][c], b = o->f[F_EY][c], d = o->f[F_EZ][c];
            const double p = o->f[F_BX][c], q = o->f[F_BY][c], r = o->f[F_BZ][c];
            e2 += a * a + b * b + d * d;
            b2 += p * p + q * q + r * r;
          }
      energy_hist[2 * n] = 0.5 * e2;
      energy_hist[2 * n + 1] = 0.5 * b2;
    }
  }
}

/* Diagnostics (no counterpart in the reference): FP64 accumulation of
 * out[0] = 1/2 sum E^2, out[1] = 1/2 sum B^2 over interior cells,
 * out[2] = max |div_h B|, out[3] = sqrt(mean (div_h B)^2) at cell centres. */
void orc_reduce(const orc_t *o, double *out) {
  double e2 = 0.0, b2 = 0.0, dmax = 0.0, d2 = 0.0;
  for (int k = 0; k < o->nz; ++k)
    for (int j = 0; j < o->ny; ++j)
      for (int i = 0; i < o->nx; ++i) {
        const int64_t c = IDX(o, i, j, k);
        const double a = o->f[F_EX][c], b = o->f[F_EY][c], d = o->f[F_EZ][c];
        const double p = o->f[F_BX][c], q = o->f[F_BY][c], r = o->f[F_BZ][c];
        e2 += a * a + b * b + d * d;
        b2 += p * p + q * q + r * r;
        const double dv = ((double)o->f[F_BX][IDX(o, i + 1, j, k)] - p) +
                          ((double)o->f[F_BY][IDX(o, i, j + 1, k)] - q) +
                          ((double)o->f[F_BZ][IDX(o, i, j, k + 1)] - r);
        if (fabs(dv) > dmax) dmax = fabs(dv);
        d2 += dv * dv;
      }
  out[0] = 0.5 * e2;
  out[1] = 0.5 * b2;
  out[2] = dmax;
  out[3] = sqrt(d2 / ((double)o->nx * o->ny * o->nz));
}

/* ------------------------------------------------------------------------
 * Output path: node interpolation and downsampling.
 *
 * Corner averages of interpFromEdges / interpFromFaces
 * (src/tools/helpers.F90:93-162, :164-257), one row per corner in the order
 * c000 c100 c010 c110 c001 c101 c011 c111, each row the (di, dj, dk) of its
 * terms in the reference's term order (the sums round in that order).
 */
static const signed char ORC_EDGE[3][8][2][3] = {
  { /* f_x, :113-126 */
    {{0, 0, 0}, {-1, 0, 0}}, {{0, 0, 0}, {1, 0, 0}}, {{0, 1, 0}, {-1, 1, 0}}, {{0, 1, 0}, {1, 1, 0}},
    {{0, 0, 1}, {-1, 0, 1}}, {{0, 0, 1}, {1, 0, 1}}, {{0, 1, 1}, {-1, 1, 1}}, {{0, 1, 1}, {1, 1, 1}} },
  { /* f_y, :130-143 */
    {{0, 0, 0}, {0, -1, 0}}, {{1, 0, 0}, {1, -1, 0}}, {{0, 0, 0}, {0, 1, 0}}, {{1, 0, 0}, {1, 1, 0}},
    {{0, 0, 1}, {0, -1, 1}}, {{1, 0, 1}, {1, -1, 1}}, {{0, 0, 1}, {0, 1, 1}}, {{1, 0, 1}, {1, 1, 1}} },
  { /* f_z, :147-154 */
    {{0, 0, 0}, {0, 0, -1}}, {{1, 0, 0}, {1, 0, -1}}, {{0, 1, 0}, {0, 1, -1}}, {{1, 1, 0}, {1, 1, -1}},
    {{0, 0, 0}, {0, 0, 1}}, {{1, 0, 0}, {1, 0, 1}}, {{0, 1, 0}, {0, 1, 1}}, {{1, 1, 0}, {1, 1, 1}} },
};
static const signed char ORC_FACE[3][8][4][3] = {
  { /* f_x, :184-199 */
    {{0, 0, 0}, {0, -1, 0}, {0, 0, -1}, {0, -1, -1}},
    {{1, 0, 0}, {1, -1, 0}, {1, 0, -1}, {1, -1, -1}},
    {{0, 0, 0}, {0, 1, 0}, {0, 0, -1}, {0, 1, -1}},
    {{1, 0, 0}, {1, 0, -1}, {1, 1, -1}, {1, 1, 0}},
    {{0, 0, 0}, {0, 0, 1}, {0, -1, 0}, {0, -1, 1}},
    {{1, 0, 0}, {1, 0, 1}, {1, -1, 0}, {1, -1, 1}},
    {{0, 0, 0}, {0, 1, 0}, {0, 1, 1}, {0, 0, 1}},
    {{1, 0, 0}, {1, 1, 0}, {1, 1, 1}, {1, 0, 1}} },
  { /* f_y, :209-224 */
    {{-1, 0, -1}, {-1, 0, 0}, {0, 0, -1}, {0, 0, 0}},
    {{0, 0, -1}, {0, 0, 0}, {1, 0, -1}, {1, 0, 0}},
    {{-1, 1, -1}, {-1, 1, 0}, {0, 1, -1}, {0, 1, 0}},
    {{0, 1, -1}, {0, 1, 0}, {1, 1, -1}, {1, 1, 0}},
    {{-1, 0, 0}, {-1, 0, 1}, {0, 0, 0}, {0, 0, 1}},
    {{0, 0, 0}, {0, 0, 1}, {1, 0, 0}, {1, 0, 1}},
    {{-1, 1, 0}, {-1, 1, 1}, {0, 1, 0}, {0, 1, 1}},
    {{0, 1, 0}, {0, 1, 1}, {1, 1, 0}, {1, 1, 1}} },
  { /* f_z, :226-241 */
    {{-1, -1, 0}, {-1, 0, 0}, {0, -1, 0}, {0, 0, 0}},
    {{0, -1, 0}, {0, 0, 0}, {1, -1, 0}, {1, 0, 0}},
    {{-1, 0, 0}, {-1, 1, 0}, {0, 0, 0}, {0, 1, 0}},
    {{0, 0, 0}, {0, 1, 0}, {1, 0, 0}, {1, 1, 0}},
    {{-1, -1, 1}, {-1, 0, 1}, {0, -1, 1}, {0, 0, 1}},
    {{0, -1, 1}, {0, 0, 1}, {1, -1, 1}, {1, 0, 1}},
    {{-1, 0, 1}, {-1, 1, 1}, {0, 0, 1}, {0, 1, 1}},
    {{0, 0, 1}, {0, 1, 1}, {1, 0, 1}, {1, 1, 1}} },
};

/* the trilinear blend of the eight corner values (every interp* routine:
 * c00 c10 c01 c11 along x, c0 c1 along y, then z) */
static real orc_trilinear(const real c[8], real dx, real dy, real dz) {
  const real c00 = c[0] * (1 - dx) + c[1] * dx;
  const real c10 = c[2] * (1 - dx) + c[3] * dx;
  const real c01 = c[4] * (1 - dx) + c[5] * dx;
  const real c11 = c[6] * (1 - dx) + c[7] * dx;
  const real c0 = c00 * (1 - dy) + c10 * dy;
  const real c1 = c01 * (1 - dy) + c11 * dy;
  return c0 * (1 - dz) + c1 * dz;
}

/* interpFromEdges (E) and interpFromFaces (B) at offset (dx, dy, dz) from
 * cell (i, j, k): out[0..2] = E, out[3..5] = B */
void orc_interp(const orc_t *o, double dx_, double dy_, double dz_, int i, int j, int k, real *out) {
  const real dx = (real)dx_, dy = (real)dy_, dz = (real)dz_;
  for (int a = 0; a < 3; ++a) {
    const real *fe = o->f[F_EX + a], *fb = o->f[F_BX + a];
    real ce[8], cb[8];
    for (int c = 0; c < 8; ++c) {
      const signed char *p = ORC_EDGE[a][c][0], *q = ORC_EDGE[a][c][1];
      ce[c] = (real)0.5 * (fe[IDX(o, i + p[0], j + p[1], k + p[2])] + fe[IDX(o, i + q[0], j + q[1], k + q[2])]);
      real s = 0;
      for (int t = 0; t < 4; ++t) {
        const signed char *r = ORC_FACE[a][c][t];
        const real v = fb[IDX(o, i + r[0], j + r[1], k + r[2])];
        s = t == 0 ? v : s + v;
      }
      cb[c] = (real)0.25 * s;
    }
    out[a] = orc_trilinear(ce, dx, dy, dz);
    out[3 + a] = orc_trilinear(cb, dx, dy, dz);
  }
}

/* Node count per axis of writeFields_hdf5's downsampled output for
 * output_istep >= 2 (src/io/write_output.F90:268-293) on one meshblock at
 * the origin: nodes 0, istep, ... below s, ceil(s / istep) of them. */
static int orc_out_n(int s, int istep) { return (s + istep - 1) / istep; }

void orc_output_dims(const orc_t *o, int istep, int *dims) {
  dims[0] = orc_out_n(o->nx, istep);
  dims[1] = orc_out_n(o->ny, istep);
  dims[2] = orc_out_n(o->nz, istep);
}

/* The six field datasets of writeFields_hdf5 (:334-372, the interpolated
 * branch): out[f] holds dims[0]*dims[1]*dims[2] values, x fastest
 * (sm_arr(i1, j1, k1)), for f = ex ey ez bx by bz. */
void orc_output_fields(const orc_t *o, int istep, real **out) {
  int d[3];
  orc_output_dims(o, istep, d);
  for (int k1 = 0; k1 < d[2]; ++k1)
    for (int j1 = 0; j1 < d[1]; ++j1)
      for (int i1 = 0; i1 < d[0]; ++i1) {
        real v[6];
        orc_interp(o, 0.0, 0.0, 0.0, i1 * istep, j1 * istep, k1 * istep, v);
        const int64_t t = i1 + (int64_t)d[0] * (j1 + (int64_t)d[1] * k1);
        for (int f = 0; f < 6; ++f) out[f][t] = v[f];
      }
}
