/* Plain-C host of the ptmf C-ABI, as a PETSc adapter would use it: create an
 * operator for an mx x my x mz Q2 box, upload coordinates, quadrature-point
 * viscosity and the Dirichlet mask, apply MatMult_MFStokes_A11 on device
 * vectors and the SpMVOp contract on host arrays, and check the error path.
 *
 * usage: abi_host mx my mz coords.bin eta.bin bc.bin x.bin y_out.bin spmv_out.bin
 * (raw little-endian doubles / bytes, layouts of include/ptmf.h)
 */
#include <stdio.h>
#include <stdlib.h>

#include "ptmf.h"

static void *slurp(const char *path, size_t bytes) {
  FILE *f = fopen(path, "rb");
  if (!f) { fprintf(stderr, "cannot open %s\n", path); exit(2); }
  void *buf = malloc(bytes);
  if (fread(buf, 1, bytes, f) != bytes) { fprintf(stderr, "short read %s\n", path); exit(2); }
  fclose(f);
  return buf;
}

static void dump(const char *path, const void *buf, size_t bytes) {
  FILE *f = fopen(path, "wb");
  if (!f || fwrite(buf, 1, bytes, f) != bytes) { fprintf(stderr, "cannot write %s\n", path); exit(2); }
  fclose(f);
}

#define CHECK(call)                                                              \
  do {                                                                           \
    int s_ = (call);                                                             \
    if (s_ != PTMF_OK) {                                                         \
      fprintf(stderr, "%s failed (%d): %s\n", #call, s_, ptmf_last_error());   \
      return 1;                                                                  \
    }                                                                            \
  } while (0)

int main(int argc, char **argv) {
  if (argc != 10) { fprintf(stderr, "usage: %s mx my mz coords eta bc x y_out spmv_out\n", argv[0]); return 2; }
  const int mx = atoi(argv[1]), my = atoi(argv[2]), mz = atoi(argv[3]);
  const long N = (long)(2 * mx + 1) * (2 * my + 1) * (2 * mz + 1), E = (long)mx * my * mz;
  double *coords = slurp(argv[4], sizeof(double) * 3 * N);
  double *eta = slurp(argv[5], sizeof(double) * 27 * E);
  unsigned char *bc = slurp(argv[6], 3 * N);
  double *x = slurp(argv[7], sizeof(double) * 3 * N);
  double *y = calloc(3 * N, sizeof(double));
  int ndev = 0;
  CHECK(ptmf_device_count(&ndev));
  if (ndev < 1) { fprintf(stderr, "no device\n"); return 3; }

  ptmf_op op = NULL;
  CHECK(ptmf_op_create(mx, my, mz, &op));
  double *dx = NULL, *dy = NULL;
  CHECK(ptmf_malloc((void **)&dx, sizeof(double) * 3 * N));
  CHECK(ptmf_malloc((void **)&dy, sizeof(double) * 3 * N));
  CHECK(ptmf_memcpy_htod(dx, x, sizeof(double) * 3 * N));
  /* MatMult before set-up is refused, as "MF operator not setup" */
  if (ptmf_a11_mult(op, dx, dy) != PTMF_ERR_NOT_SETUP) { fprintf(stderr, "expected NOT_SETUP\n"); return 1; }
  CHECK(ptmf_op_set_coordinates(op, coords));
  CHECK(ptmf_op_set_viscosity(op, eta));
  CHECK(ptmf_op_set_dirichlet(op, bc));
  CHECK(ptmf_a11_mult(op, dx, dy));
  CHECK(ptmf_memcpy_dtoh(y, dy, sizeof(double) * 3 * N));
  dump(argv[8], y, sizeof(double) * 3 * N);
  /* SpMVOp_MatMult contract on host arrays: Yu = A11_local u (the caller zeroes Yu first) */
  for (long i = 0; i < 3 * N; i++) y[i] = 0.0;
  CHECK(ptmf_a11_spmv_host(op, x, y));
  dump(argv[9], y, sizeof(double) * 3 * N);
  CHECK(ptmf_free(dx));
  CHECK(ptmf_free(dy));
  CHECK(ptmf_op_destroy(&op));
  if (op != NULL) { fprintf(stderr, "destroy did not clear the handle\n"); return 1; }
  printf("ok\n");
  free(coords); free(eta); free(bc); free(x); free(y);
  return 0;
}
