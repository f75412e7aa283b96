/* tff_io.c -- libtff_io.so: the output files of writeOutput
 * (src/io/write_output.F90:27-500) through the serial HDF5 C API, declared
 * in include/tff_io.h.  Host C (no C++ runtime: the HDF5 library's
 * directory also holds an older libstdc++ than the HIP runtime needs, and
 * this library must not pull it in).  The field datasets come from
 * libtff.so (tff_output_fields / tff_output_samples). */
#include "tff_io.h"

#include <errno.h>
#include <hdf5.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>

static _Thread_local char io_err[4608];

static int io_fail(int code, const char* msg, const char* detail) {
  snprintf(io_err, sizeof io_err, "%s%s%s", msg, detail ? ": " : "", detail ? detail : "");
  return code;
}

const char* tff_io_last_error(void) { return io_err; }

/* "<dir>/<stem>.NNNNN" (write_output.F90:73-74, i5.5) */
static void numbered(char* out, size_t cap, const char* dir, const char* stem, int index, const char* suffix) {
  snprintf(out, cap, "%s/%s.%05d%s", dir, stem, index, suffix);
}

/* one dataset of `rank` dims (C order) of memory type `mt`, written whole */
static int write_dataset(hid_t file, const char* name, hid_t mt, int rank, const hsize_t* dims, const void* data) {
  const hid_t space = H5Screate_simple(rank, dims, NULL);
  if (space < 0) return 0;
  const hid_t ds = H5Dcreate2(file, name, mt, space, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
  const int ok = ds >= 0 && H5Dwrite(ds, mt, H5S_ALL, H5S_ALL, H5P_DEFAULT, data) >= 0;
  if (ds >= 0) H5Dclose(ds);
  H5Sclose(space);
  return ok;
}

static const char* const FLD_VARS[12] = {"ex", "ey", "ez", "bx", "by", "bz",
                                         "var1", "var2", "var3", "xx", "yy", "zz"};

/* writeXDMF_hdf5 (write_output.F90:116-182): the same lines, the three
 * dimensions in Fortran's I10 as (nk, nj, ni) */
static int write_xdmf(const char* dir, int index, const int dims[3]) {
  char path[4096];
  numbered(path, sizeof path, dir, "flds.tot", index, ".xdmf");
  FILE* f = fopen(path, "w");
  if (!f) return 0;
  fprintf(f, "<?xml version=\"1.0\" ?>\n");
  fprintf(f, "<!DOCTYPE Xdmf SYSTEM \"Xdmf.dtd\">\n");
  fprintf(f, "<Xdmf Version=\"2.0\">\n");
  fprintf(f, "  <Domain>\n");
  fprintf(f, "    <Grid Name=\"domain\" GridType=\"Uniform\">\n");
  fprintf(f, "      <Topology TopologyType=\"3DCoRectMesh\" Dimensions=\"%10d%10d%10d\"/>\n", dims[2], dims[1],
          dims[0]);
  fprintf(f, "      <Geometry GeometryType=\"ORIGIN_DXDYDZ\">\n");
  fprintf(f, "        <DataItem Format=\"XML\" Dimensions=\"3\" NumberType=\"Float\" Precision=\"4\">\n");
  fprintf(f, "          0.0 0.0 0.0\n");
  fprintf(f, "        </DataItem>\n");
  fprintf(f, "        <DataItem Format=\"XML\" Dimensions=\"3\" NumberType=\"Float\" Precision=\"4\">\n");
  fprintf(f, "          1.0 1.0 1.0\n");
  fprintf(f, "        </DataItem>\n");
  fprintf(f, "      </Geometry>\n");
  for (int v = 0; v < 12; ++v) {
    fprintf(f, "      <Attribute Name=\"%s\" Center=\"Node\">\n", FLD_VARS[v]);
    fprintf(f, "        <DataItem Format=\"HDF\" Dimensions=\"%10d%10d%10d\" NumberType=\"Float\" Precision=\"4\">\n",
            dims[2], dims[1], dims[0]);
    fprintf(f, "          flds.tot.%05d:/%s\n", index, FLD_VARS[v]);
    fprintf(f, "        </DataItem>\n");
    fprintf(f, "      </Attribute>\n");
  }
  fprintf(f, "    </Grid>\n");
  fprintf(f, "  </Domain>\n");
  fprintf(f, "</Xdmf>\n");
  return fclose(f) == 0;
}

/* mkdir -p (initialize.F90:242-254 creates the output directory) */
static int make_dir(const char* dir) {
  char p[4096];
  const size_t n = strlen(dir);
  if (n == 0 || n >= sizeof p) return 0;
  memcpy(p, dir, n + 1);
  for (size_t i = 1; i <= n; ++i) {
    if (p[i] == '/' || p[i] == '\0') {
      const char c = p[i];
      p[i] = '\0';
      if (mkdir(p, 0777) != 0 && errno != EEXIST) return 0;
      p[i] = c;
    }
  }
  struct stat st;
  return stat(dir, &st) == 0 && S_ISDIR(st.st_mode);
}

int tff_io_write_params(const char* dir, int index, int timestep, int n, const char* names, const int* types,
                        const int* ivals, const float* rvals) {
  if (!dir || n < 0 || (n > 0 && (!names || !types || !ivals || !rvals)))
    return io_fail(TFF_EINVAL, "tff_io_write_params: bad arguments", NULL);
  char path[4096];
  numbered(path, sizeof path, dir, "params", index, "");
  const hid_t file = H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
  if (file < 0) return io_fail(TFF_EINVAL, "cannot create", path);
  const hsize_t one = 1;
  int rc = TFF_OK;
  if (!write_dataset(file, "timestep", H5T_NATIVE_INT, 1, &one, &timestep))
    rc = io_fail(TFF_EINVAL, "params: dataset", "timestep");
  const char* name = names;
  for (int q = 0; q < n && rc == TFF_OK; ++q) {
    int ok;
    if (types[q] == 1 || types[q] == 3)
      ok = write_dataset(file, name, H5T_NATIVE_INT, 1, &one, &ivals[q]);
    else if (types[q] == 2)
      ok = write_dataset(file, name, H5T_NATIVE_FLOAT, 1, &one, &rvals[q]);
    else
      ok = 0;  /* write_output.F90:106: unknown param_type */
    if (!ok) rc = io_fail(TFF_EINVAL, "params: dataset", name);
    name += strlen(name) + 1;
  }
  H5Fclose(file);
  return rc;
}

int tff_io_write_fields(const char* dir, int index, int istep, const int dims[3], const float* const data[9]) {
  if (!dir || !dims || !data || istep < 1 || dims[0] < 1 || dims[1] < 1 || dims[2] < 1)
    return io_fail(TFF_EINVAL, "tff_io_write_fields: bad arguments", NULL);
  for (int f = 0; f < 9; ++f)
    if (!data[f]) return io_fail(TFF_EINVAL, "tff_io_write_fields: null dataset", FLD_VARS[f]);
  if (!write_xdmf(dir, index, dims)) return io_fail(TFF_EINVAL, "cannot write the XDMF wrapper in", dir);
  char path[4096];
  numbered(path, sizeof path, dir, "flds.tot", index, "");
  const hid_t file = H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
  if (file < 0) return io_fail(TFF_EINVAL, "cannot create", path);
  const hsize_t hd[3] = {(hsize_t)dims[2], (hsize_t)dims[1], (hsize_t)dims[0]};
  int rc = TFF_OK;
  for (int f = 0; f < 9 && rc == TFF_OK; ++f)
    if (!write_dataset(file, FLD_VARS[f], H5T_NATIVE_FLOAT, 3, hd, data[f]))
      rc = io_fail(TFF_EINVAL, "flds.tot: dataset", FLD_VARS[f]);
  /* xx, yy, zz: REAL(x0 + i) at the written nodes (write_output.F90:384-389) */
  const size_t nout = (size_t)dims[0] * dims[1] * dims[2];
  float* c = rc == TFF_OK ? (float*)malloc(nout * sizeof(float)) : NULL;
  if (rc == TFF_OK && !c) rc = io_fail(TFF_ENOMEM, "flds.tot: host memory for the coordinates", NULL);
  for (int a = 0; a < 3 && rc == TFF_OK; ++a) {
    size_t t = 0;
    for (int k = 0; k < dims[2]; ++k)
      for (int j = 0; j < dims[1]; ++j)
        for (int i = 0; i < dims[0]; ++i) c[t++] = (float)((a == 0 ? i : (a == 1 ? j : k)) * istep);
    if (!write_dataset(file, FLD_VARS[9 + a], H5T_NATIVE_FLOAT, 3, hd, c))
      rc = io_fail(TFF_EINVAL, "flds.tot: dataset", FLD_VARS[9 + a]);
  }
  free(c);
  H5Fclose(file);
  return rc;
}

int tff_io_write_domain(const char* dir, int index, int n, const int* blocks) {
  if (!dir || n < 1 || !blocks) return io_fail(TFF_EINVAL, "tff_io_write_domain: bad arguments", NULL);
  static const char* const vars[6] = {"x0", "y0", "z0", "sx", "sy", "sz"};
  char path[4096];
  numbered(path, sizeof path, dir, "domain", index, "");
  const hid_t file = H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
  if (file < 0) return io_fail(TFF_EINVAL, "cannot create", path);
  int* col = (int*)malloc((size_t)n * sizeof(int));
  int rc = col ? TFF_OK : io_fail(TFF_ENOMEM, "domain: host memory", NULL);
  const hsize_t len = (hsize_t)n;
  for (int d = 0; d < 6 && rc == TFF_OK; ++d) {
    for (int r = 0; r < n; ++r) col[r] = blocks[6 * r + d];
    if (!write_dataset(file, vars[d], H5T_NATIVE_INT, 1, &len, col)) rc = io_fail(TFF_EINVAL, "domain: dataset", vars[d]);
  }
  free(col);
  H5Fclose(file);
  return rc;
}

int tff_io_write_output(tff_handle* h, const char* dir, int index, int timestep, int istep, int user_output, int n,
                        const char* names, const int* types, const int* ivals, const float* rvals) {
  if (!h || !dir) return io_fail(TFF_EINVAL, "tff_io_write_output: bad arguments", NULL);
  if (istep < 2)
    return io_fail(TFF_EINVAL, "output istep must be >= 2 (the istep = 1 debug layout with ghosts is not supported)",
                   NULL);
  if (user_output < TFF_USEROUT_NONE || user_output > TFF_USEROUT_UNIFORM)
    return io_fail(TFF_EINVAL, "tff_io_write_output: unknown user output kind", NULL);
  if (!make_dir(dir)) return io_fail(TFF_EINVAL, "cannot create the output directory", dir);
  int box[3], dims[3];
  int rc = tff_get_size(h, box, NULL);
  if (rc == TFF_OK) rc = tff_output_dims(h, istep, dims);
  if (rc) return io_fail(rc, "libtff", tff_last_error(h));
  const size_t nout = (size_t)dims[0] * dims[1] * dims[2];
  float* buf = (float*)malloc(nout * 16 * sizeof(float));
  if (!buf) return io_fail(TFF_ENOMEM, "tff_io_write_output: host memory", NULL);
  float* d[16];
  for (int f = 0; f < 16; ++f) d[f] = buf + (size_t)f * nout;
  void* fo[6] = {d[0], d[1], d[2], d[3], d[4], d[5]};
  rc = tff_output_fields(h, istep, fo, TFF_HOST_F32);
  if (rc) rc = io_fail(rc, "libtff", tff_last_error(h));
  /* var1..var3 (write_output.F90:375-383) */
  float* var[3] = {d[6], d[7], d[8]};
  for (int v = 0; v < 3; ++v) memset(var[v], 0, nout * sizeof(float));
  if (rc == TFF_OK && user_output != TFF_USEROUT_NONE) {
    /* samples: raw ex ey ez bx by bz and rho at the nodes */
    float* s[7] = {d[9], d[10], d[11], d[12], d[13], d[14], d[15]};
    void* so[7] = {s[0], s[1], s[2], s[3], s[4], s[5], s[6]};
    rc = tff_output_samples(h, istep, so, TFF_HOST_F32);
    if (rc) rc = io_fail(rc, "libtff", tff_last_error(h));
    if (rc == TFF_OK) {
      memcpy(var[0], s[6], nout * sizeof(float));
      if (user_output == TFF_USEROUT_UNIFORM) {
        /* user_uniform.F90:159-169: |(-bz ex + bx ez, bz ey - by ez)| over
         * the distance from the box axis and over |B|^2, in default REAL */
        const float tiny = 1e-7f, hx = 0.5f * (float)box[0], hy = 0.5f * (float)box[1];
        size_t t = 0;
        for (int k = 0; k < dims[2]; ++k)
          for (int j = 0; j < dims[1]; ++j)
            for (int i = 0; i < dims[0]; ++i, ++t) {
              const float ex = s[0][t], ey = s[1][t], ez = s[2][t], bx = s[3][t], by = s[4][t], bz = s[5][t];
              const float a = -bz * ex + bx * ez, b = bz * ey - by * ez;
              float temp = sqrtf(a * a + b * b);
              const float d1 = (float)(i * istep) - hx, d2 = (float)(j * istep) - hy;
              temp = temp / sqrtf(d1 * d1 + d2 * d2 + tiny);
              temp = temp / (bx * bx + by * by + bz * bz);
              var[1][t] = temp;
            }
      }
    }
  }
  if (rc == TFF_OK) rc = tff_io_write_params(dir, index, timestep, n, names, types, ivals, rvals);
  if (rc == TFF_OK) {
    const float* data[9] = {d[0], d[1], d[2], d[3], d[4], d[5], d[6], d[7], d[8]};
    rc = tff_io_write_fields(dir, index, istep, dims, data);
  }
  if (rc == TFF_OK) {
    const int blocks[6] = {0, 0, 0, box[0], box[1], box[2]};
    rc = tff_io_write_domain(dir, index, 1, blocks);
  }
  free(buf);
  return rc;
}
