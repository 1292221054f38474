/* h5_interop.c — test helper linked against a real libhdf5.
 *
 *   h5_interop write <file>          creates a file the way the reference's
 *       Mesh::Save and myhdfshell.cpp create theirs (H5Fcreate, H5Screate_simple,
 *       H5Dcreate with H5T_STD_I32LE / H5T_IEEE_F64LE / H5T_STD_I64LE and the
 *       SavedIon compound of save.cpp:256-281, H5Dwrite of native data), with
 *       values given by closed forms the test recomputes.
 *   h5_interop f64 <file> <name>     prints the dataset read as native doubles
 *   h5_interop i64 <file> <name>     ... as native 64-bit integers
 *   h5_interop ion <file> <name>     ... as SavedIon records, one per line
 * Every read prints "dims d0 d1 ..." first; values use %.17g.  Exit status 0
 * on success, 1 on any library error. */
#include <hdf5.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
  int cell;
  double x, y, z, px, py, pz, weight, q2m;
  int i_z, sort;
} Ion;

static hid_t ion_type(void) {
  hid_t t = H5Tcreate(H5T_COMPOUND, sizeof(Ion));
  H5Tinsert(t, "cell", HOFFSET(Ion, cell), H5T_NATIVE_INT);
  H5Tinsert(t, "x", HOFFSET(Ion, x), H5T_IEEE_F64LE);
  H5Tinsert(t, "y", HOFFSET(Ion, y), H5T_IEEE_F64LE);
  H5Tinsert(t, "z", HOFFSET(Ion, z), H5T_IEEE_F64LE);
  H5Tinsert(t, "px", HOFFSET(Ion, px), H5T_IEEE_F64LE);
  H5Tinsert(t, "py", HOFFSET(Ion, py), H5T_IEEE_F64LE);
  H5Tinsert(t, "pz", HOFFSET(Ion, pz), H5T_IEEE_F64LE);
  H5Tinsert(t, "weight", HOFFSET(Ion, weight), H5T_IEEE_F64LE);
  H5Tinsert(t, "q2m", HOFFSET(Ion, q2m), H5T_IEEE_F64LE);
  H5Tinsert(t, "i_z", HOFFSET(Ion, i_z), H5T_NATIVE_INT);
  H5Tinsert(t, "sort", HOFFSET(Ion, sort), H5T_NATIVE_INT);
  return t;
}

static int put(hid_t f, const char *name, hid_t ftype, hid_t mtype, int rank,
               const hsize_t *dims, const void *data) {
  hid_t s = H5Screate_simple(rank, dims, NULL);
  hid_t d = H5Dcreate(f, name, ftype, s, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
  if (s < 0 || d < 0) return 1;
  herr_t e = H5Dwrite(d, mtype, H5S_ALL, H5S_ALL, H5P_DEFAULT, data);
  H5Dclose(d);
  H5Sclose(s);
  return e < 0;
}

static int do_write(const char *path) {
  hid_t f = H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
  if (f < 0) return 1;
  int err = 0;
  int mi[16];
  for (int i = 0; i < 16; ++i) mi[i] = 3 * i - 7;
  hsize_t d16 = 16, d1 = 1;
  err |= put(f, "meshinfo", H5T_STD_I32LE, H5T_NATIVE_INT, 1, &d16, mi);
  double phase = 2.75;
  err |= put(f, "phase", H5T_IEEE_F64LE, H5T_NATIVE_DOUBLE, 1, &d1, &phase);
  /* 3-D fields: v(k, j, i) = i + 10 j + 100 k + 0.25 f  over [5][4][6] */
  const int NZ = 5, NY = 4, NX = 6;
  hsize_t d3[3] = {NZ, NY, NX};
  double *v = malloc(sizeof(double) * NZ * NY * NX);
  const char *fields[] = {"ex", "ey", "ez", "bx", "by", "bz", "jx", "jy", "jz", "dens", "Dens0", "Dens1"};
  for (int fi = 0; fi < 12; ++fi) {
    for (int k = 0; k < NZ; ++k)
      for (int j = 0; j < NY; ++j)
        for (int i = 0; i < NX; ++i) v[i + NX * (j + NY * k)] = i + 10. * j + 100. * k + 0.25 * fi;
    err |= put(f, fields[fi], H5T_IEEE_F64LE, H5T_NATIVE_DOUBLE, 3, d3, v);
  }
  free(v);
  /* particles: record r has cell 1000 + r, x = r + 0.5, ..., sort 1 */
  enum { NP = 37 };
  Ion p[NP];
  memset(p, 0, sizeof p);
  for (int r = 0; r < NP; ++r) {
    p[r].cell = 1000 + r;
    p[r].x = r + 0.5; p[r].y = -r; p[r].z = 0.125 * r;
    p[r].px = 1e-3 * r; p[r].py = 2.0; p[r].pz = -3.0;
    p[r].weight = 1.0 / (r + 1); p[r].q2m = -1.0;
    p[r].i_z = 1; p[r].sort = 1;
  }
  hsize_t dp = NP;
  hid_t it = ion_type();
  err |= put(f, "Ions_of_sort_01_at_PE_0", it, it, 1, &dp, p);
  H5Tclose(it);
  long long np[2] = {0, NP};
  hsize_t d2 = 2;
  err |= put(f, "Nparticles_at_PE_0", H5T_STD_I64LE, H5T_NATIVE_LLONG, 1, &d2, np);
  err |= H5Fclose(f) < 0;
  return err;
}

static int do_read(const char *path, const char *name, const char *what) {
  hid_t f = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
  if (f < 0) return 1;
  hid_t d = H5Dopen(f, name, H5P_DEFAULT);
  if (d < 0) return 1;
  hid_t s = H5Dget_space(d);
  int rank = H5Sget_simple_extent_ndims(s);
  hsize_t dims[8];
  H5Sget_simple_extent_dims(s, dims, NULL);
  size_t n = 1;
  printf("dims");
  for (int r = 0; r < rank; ++r) {
    printf(" %llu", (unsigned long long)dims[r]);
    n *= dims[r];
  }
  printf("\n");
  int err = 0;
  if (!strcmp(what, "f64")) {
    double *v = malloc(sizeof(double) * (n ? n : 1));
    err = H5Dread(d, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, v) < 0;
    for (size_t i = 0; i < n && !err; ++i) printf("%.17g\n", v[i]);
    free(v);
  } else if (!strcmp(what, "i64")) {
    long long *v = malloc(sizeof(long long) * (n ? n : 1));
    err = H5Dread(d, H5T_NATIVE_LLONG, H5S_ALL, H5S_ALL, H5P_DEFAULT, v) < 0;
    for (size_t i = 0; i < n && !err; ++i) printf("%lld\n", v[i]);
    free(v);
  } else {
    Ion *v = malloc(sizeof(Ion) * (n ? n : 1));
    hid_t it = ion_type();
    err = H5Dread(d, it, H5S_ALL, H5S_ALL, H5P_DEFAULT, v) < 0;
    for (size_t i = 0; i < n && !err; ++i)
      printf("%d %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %d %d\n", v[i].cell, v[i].x,
             v[i].y, v[i].z, v[i].px, v[i].py, v[i].pz, v[i].weight, v[i].q2m, v[i].i_z,
             v[i].sort);
    H5Tclose(it);
    free(v);
  }
  H5Sclose(s);
  H5Dclose(d);
  H5Fclose(f);
  return err;
}

int main(int argc, char **argv) {
  if (argc == 3 && !strcmp(argv[1], "write")) return do_write(argv[2]);
  if (argc == 4) return do_read(argv[2], argv[3], argv[1]);
  fprintf(stderr, "usage: h5_interop write <file> | {f64,i64,ion} <file> <name>\n");
  return 2;
}
