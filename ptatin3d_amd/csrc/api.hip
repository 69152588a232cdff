// api.hip -- C-ABI entry points of the ptmf operator library (see include/ptmf.h).
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <vector>

#include "ptmf_internal.h"

namespace {
thread_local char g_err[1024] = "";
}

int ptmf_set_error(int code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
  return code;
}

using ptmf::Box;

namespace ptmf {

Tiling make_tiling(int mx, int my, int mz, int TX, int TY, int ZC) {
  Tiling t;
  t.TX = TX; t.TY = TY; t.ZC = ZC;
  t.ntx = (mx + TX - 1) / TX;
  t.nty = (my + TY - 1) / TY;
  t.ntz = (mz + ZC - 1) / ZC;
  t.ntiles = (long)t.ntx * t.nty * t.ntz;
  return t;
}

Tiling choose_tiling(int mx, int my, int mz) {
  int bestx = 1, besty = 1;
  double best = -1.0;
  for (int TX = 1; TX <= 21; TX++)
    for (int TY = 1; TX * TY <= 21; TY++) {
      if ((2 * TX + 1) * (2 * TY + 1) > 130) continue;
      const long tiles = (long)((mx + TX - 1) / TX) * ((my + TY - 1) / TY);
      const double fill = (double)mx * my / (tiles * 21.0);
      const double face = 1.0 / TX + 1.0 / TY;  // face-buffer traffic grows with this
      const double score = fill - 0.02 * face;
      if (score > best + 1e-12) {
        best = score; bestx = TX; besty = TY;
      }
    }
  const long txy = (long)((mx + bestx - 1) / bestx) * ((my + besty - 1) / besty);
  // about 6000 waves per launch (measured best at 64^3: z-chunks of 2); the
  // small multigrid levels get z-chunks of 1 to keep every SIMD busy
  long zc = (long)mz * txy / 6000;
  if (zc < 1) zc = 1;
  if (zc > 16) zc = 16;
  if (zc > mz) zc = mz;
  return make_tiling(mx, my, mz, bestx, besty, (int)zc);
}

}  // namespace ptmf

extern "C" {

const char *ptmf_last_error(void) { return g_err; }

int ptmf_version(int *major, int *minor) {
  if (major) *major = 0;
  if (minor) *minor = 1;
  return PTMF_OK;
}

int ptmf_device_count(int *n) {
  PTMF_CHECK(n, PTMF_ERR_ARG, "device_count: null");
  int c = 0;
  hipError_t e = hipGetDeviceCount(&c);
  if (e != hipSuccess) c = 0;
  *n = c;
  return PTMF_OK;
}

int ptmf_set_device(int dev) {
  PTMF_HIP(hipSetDevice(dev));
  return PTMF_OK;
}

int ptmf_malloc(void **dptr, size_t bytes) {
  PTMF_CHECK(dptr, PTMF_ERR_ARG, "malloc: null");
  *dptr = nullptr;
  if (bytes == 0) return PTMF_OK;
  hipError_t e = hipMalloc(dptr, bytes);
  if (e != hipSuccess) {
    *dptr = nullptr;
    return ptmf_set_error(PTMF_ERR_NOMEM, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
  }
  return PTMF_OK;
}

int ptmf_free(void *dptr) {
  if (dptr) PTMF_HIP(hipFree(dptr));
  return PTMF_OK;
}

// The helpers below first wait for all work on the device, so they are
// ordered after asynchronous entry points on any operator's stream.
int ptmf_memcpy_htod(void *dst, const void *src, size_t bytes) {
  if (!bytes) return PTMF_OK;
  PTMF_HIP(hipDeviceSynchronize());
  PTMF_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  return PTMF_OK;
}

int ptmf_memcpy_dtoh(void *dst, const void *src, size_t bytes) {
  if (!bytes) return PTMF_OK;
  PTMF_HIP(hipDeviceSynchronize());
  PTMF_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  return PTMF_OK;
}

int ptmf_memcpy_dtod(void *dst, const void *src, size_t bytes) {
  PTMF_CHECK(bytes % 16 == 0 && ((uintptr_t)dst % 16) == 0 && ((uintptr_t)src % 16) == 0, PTMF_ERR_ARG,
             "memcpy_dtod: 16-byte aligned pointers and sizes only");
  PTMF_HIP(hipDeviceSynchronize());
  PTMF_TRY(ptmf::launch_copy16(src, dst, (long)(bytes / 16), 0));
  PTMF_HIP(hipStreamSynchronize(0));
  return PTMF_OK;
}

int ptmf_memset(void *dptr, int value, size_t bytes) {
  if (!bytes) return PTMF_OK;
  PTMF_HIP(hipDeviceSynchronize());
  PTMF_HIP(hipMemset(dptr, value, bytes));
  return PTMF_OK;
}

int ptmf_device_synchronize(void) {
  PTMF_HIP(hipDeviceSynchronize());
  return PTMF_OK;
}

// ---------------------------------------------------------------------------
// operator lifecycle
// ---------------------------------------------------------------------------
int ptmf_op_create(int mx, int my, int mz, ptmf_op *out) {
  PTMF_CHECK(out, PTMF_ERR_ARG, "op_create: null handle");
  *out = nullptr;
  PTMF_CHECK(mx > 0 && my > 0 && mz > 0, PTMF_ERR_ARG, "op_create: bad box %d x %d x %d", mx, my, mz);
  ptmf_op_s *op = new ptmf_op_s();
  op->box = ptmf::make_box(mx, my, mz);
  const Box &b = op->box;
  int st = PTMF_OK;
  auto alloc = [&](void **p, size_t n) {
    if (st == PTMF_OK) st = ptmf_malloc(p, n);
  };
  alloc((void **)&op->d64.coords, sizeof(double) * 3 * b.nnodes);
  alloc((void **)&op->d64.eta, sizeof(double) * 27 * b.nel);
  alloc((void **)&op->d32.coords, sizeof(float) * 3 * b.nnodes);
  alloc((void **)&op->d32.eta, sizeof(float) * 27 * b.nel);
  alloc((void **)&op->bcnode, sizeof(uint8_t) * b.nnodes);
  if (st == PTMF_OK) {
    hipError_t e = hipGetDevice(&op->device);
    if (e == hipSuccess) e = hipStreamCreate(&op->stream);
    if (e == hipSuccess) e = hipEventCreate(&op->ev0);
    if (e == hipSuccess) e = hipEventCreate(&op->ev1);
    if (e != hipSuccess) st = ptmf_set_error(PTMF_ERR_DEVICE, "op_create: %s", hipGetErrorString(e));
  }
  if (st == PTMF_OK) st = ptmf_memset(op->bcnode, 0, b.nnodes);
  if (st == PTMF_OK) {
    op->tiling = ptmf::choose_tiling(mx, my, mz);
    op->faces_doubles = ptmf::face_buffer_doubles(op->tiling.TX, op->tiling.TY, op->tiling.ZC,
                                                  op->tiling.ntiles);
    st = ptmf_malloc(&op->faces, sizeof(double) * op->faces_doubles);
  }
  if (st != PTMF_OK) {
    ptmf_op tmp = op;
    ptmf_op_destroy(&tmp);
    return st;
  }
  *out = op;
  return PTMF_OK;
}

int ptmf_op_destroy(ptmf_op *pop) {
  if (!pop || !*pop) return PTMF_OK;
  ptmf_op_s *op = *pop;
  if (op->stream) (void)hipStreamSynchronize(op->stream);
  (void)hipFree(op->d64.coords);
  (void)hipFree(op->d64.eta);
  (void)hipFree(op->d32.coords);
  (void)hipFree(op->d32.eta);
  (void)hipFree(op->bcnode);
  (void)hipFree(op->bcidx);
  (void)hipFree(op->faces);
  (void)hipFree(op->host_stage);
  (void)hipFree(op->host_stage2);
  (void)hipFree(op->pa64);
  (void)hipFree(op->pa32);
  (void)hipFree(op->pa_bad);
  (void)hipFree(op->et64);
  (void)hipFree(op->et32);
  if (op->ev0) (void)hipEventDestroy(op->ev0);
  if (op->ev1) (void)hipEventDestroy(op->ev1);
  if (op->stream) (void)hipStreamDestroy(op->stream);
  delete op;
  *pop = nullptr;
  return PTMF_OK;
}

int ptmf_op_sizes(ptmf_op op, long *nnodes, long *nel) {
  PTMF_CHECK(op, PTMF_ERR_ARG, "op_sizes: null op");
  if (nnodes) *nnodes = op->box.nnodes;
  if (nel) *nel = op->box.nel;
  return PTMF_OK;
}

int ptmf_op_set_tiling(ptmf_op op, int TX, int TY, int ZC) {
  PTMF_CHECK(op, PTMF_ERR_ARG, "set_tiling: null op");
  PTMF_CHECK(TX > 0 && TY > 0 && ZC > 0 && TX * TY <= 21 && (2 * TX + 1) * (2 * TY + 1) <= 130,
             PTMF_ERR_ARG, "set_tiling: unsupported tile %d x %d x %d (TX*TY <= 21)", TX, TY, ZC);
  const Box &b = op->box;
  ptmf::Tiling t = ptmf::make_tiling(b.mx, b.my, b.mz, TX, TY < b.my ? TY : TY, ZC);
  const long need = ptmf::face_buffer_doubles(TX, TY, ZC, t.ntiles);
  PTMF_HIP(hipStreamSynchronize(op->stream));
  if (need > op->faces_doubles) {
    void *p = nullptr;
    PTMF_TRY(ptmf_malloc(&p, sizeof(double) * need));
    PTMF_HIP(hipFree(op->faces));
    op->faces = p;
    op->faces_doubles = need;
  }
  op->tiling = t;
  return op->has_eta ? ptmf::prepare_eta_tiled(op) : PTMF_OK;
}

int ptmf_op_set_kernel_variant(ptmf_op op, int variant) {
  PTMF_CHECK(op, PTMF_ERR_ARG, "set_kernel_variant: null operator");
  PTMF_CHECK(variant >= 0 && variant <= 3, PTMF_ERR_UNSUPPORTED,
             "set_kernel_variant: unknown variant %d", variant);
  op->variant = variant;
  return PTMF_OK;
}

int ptmf_op_set_tiles_per_group(ptmf_op op, int tiles) {
  PTMF_CHECK(op, PTMF_ERR_ARG, "set_tiles_per_group: null operator");
  PTMF_CHECK(tiles == 1 || tiles == 2 || tiles == 4, PTMF_ERR_UNSUPPORTED,
             "set_tiles_per_group: %d (1, 2 or 4)", tiles);
  PTMF_HIP(hipStreamSynchronize(op->stream));
  op->wg_tiles = tiles;
  return PTMF_OK;
}

int ptmf_op_get_tiles_per_group(ptmf_op op, int *tiles) {
  PTMF_CHECK(op && tiles, PTMF_ERR_ARG, "get_tiles_per_group: null argument");
  *tiles = op->wg_tiles;
  return PTMF_OK;
}

int ptmf_op_get_kernel_variant(ptmf_op op, int *variant) {
  PTMF_CHECK(op && variant, PTMF_ERR_ARG, "get_kernel_variant: null argument");
  *variant = op->variant;
  return PTMF_OK;
}

int ptmf_op_get_tiling(ptmf_op op, int *TX, int *TY, int *ZC) {
  PTMF_CHECK(op, PTMF_ERR_ARG, "get_tiling: null op");
  if (TX) *TX = op->tiling.TX;
  if (TY) *TY = op->tiling.TY;
  if (ZC) *ZC = op->tiling.ZC;
  return PTMF_OK;
}

int ptmf_op_state(ptmf_op op, unsigned long *state) {
  PTMF_CHECK(op && state, PTMF_ERR_ARG, "op_state: null");
  *state = op->state;
  return PTMF_OK;
}

int ptmf_op_set_coordinates(ptmf_op op, const double *coords) {
  PTMF_CHECK(op && coords, PTMF_ERR_ARG, "set_coordinates: null");
  const long n = 3 * op->box.nnodes;
  std::vector<float> c32(n);
  for (long i = 0; i < n; i++) c32[i] = (float)coords[i];
  PTMF_HIP(hipMemcpy(op->d64.coords, coords, sizeof(double) * n, hipMemcpyHostToDevice));
  PTMF_HIP(hipMemcpy(op->d32.coords, c32.data(), sizeof(float) * n, hipMemcpyHostToDevice));
  op->has_coords = 1;
  op->state++;
  return PTMF_OK;
}

int ptmf_op_set_coordinates_device(ptmf_op op, const double *coords) {
  PTMF_CHECK(op && coords, PTMF_ERR_ARG, "set_coordinates_device: null");
  const long n = 3 * op->box.nnodes;
  std::vector<double> h(n);
  PTMF_HIP(hipMemcpy(h.data(), coords, sizeof(double) * n, hipMemcpyDeviceToHost));
  return ptmf_op_set_coordinates(op, h.data());
}

// viscosity arrives element-major (eta[27e+q]) and is stored qp-major
// (eta[q*nel+e]) so the lanes of a wave read consecutive addresses.
int ptmf_op_set_viscosity(ptmf_op op, const double *eta) {
  PTMF_CHECK(op && eta, PTMF_ERR_ARG, "set_viscosity: null");
  const long nel = op->box.nel;
  std::vector<double> t64(27 * nel);
  std::vector<float> t32(27 * nel);
  for (long e = 0; e < nel; e++)
    for (int q = 0; q < 27; q++) {
      const double v = eta[27 * e + q];
      t64[(long)q * nel + e] = v;
      t32[(long)q * nel + e] = (float)v;
    }
  PTMF_HIP(hipMemcpy(op->d64.eta, t64.data(), sizeof(double) * 27 * nel, hipMemcpyHostToDevice));
  PTMF_HIP(hipMemcpy(op->d32.eta, t32.data(), sizeof(float) * 27 * nel, hipMemcpyHostToDevice));
  op->has_eta = 1;
  op->state++;
  // the MatMult's tile-ordered copy is built here, at set-up, so that the
  // allocation-free MatMult finds it allocated
  return ptmf::prepare_eta_tiled(op);
}

int ptmf_op_set_viscosity_device(ptmf_op op, const double *eta) {
  PTMF_CHECK(op && eta, PTMF_ERR_ARG, "set_viscosity_device: null");
  PTMF_TRY(ptmf::launch_eta_qpmajor(op->box.nel, eta, op->d64.eta, op->d32.eta, op->stream));
  op->has_eta = 1;
  op->state++;
  PTMF_TRY(ptmf::prepare_eta_tiled(op));
  PTMF_HIP(hipStreamSynchronize(op->stream));
  return PTMF_OK;
}

int ptmf_op_powerlaw_viscosity(ptmf_op op, const double *u, const double *B, double nexp, double eta_min,
                               double eta_max, double *eta) {
  PTMF_CHECK(op && u && B && eta, PTMF_ERR_ARG, "powerlaw_viscosity: null");
  PTMF_CHECK(op->has_coords, PTMF_ERR_NOT_SETUP, "powerlaw_viscosity: coordinates not set");
  PTMF_CHECK(nexp > 0.0 && eta_min > 0.0 && eta_max >= eta_min, PTMF_ERR_ARG,
             "powerlaw_viscosity: n %g, bounds [%g, %g]", nexp, eta_min, eta_max);
  PTMF_TRY(ptmf::launch_powerlaw(op, u, B, nexp, eta_min, eta_max, eta));
  PTMF_HIP(hipStreamSynchronize(op->stream));
  return PTMF_OK;
}

int ptmf_op_set_dirichlet(ptmf_op op, const unsigned char *mask) {
  PTMF_CHECK(op && mask, PTMF_ERR_ARG, "set_dirichlet: null");
  const Box &b = op->box;
  std::vector<uint8_t> bits(b.nnodes, 0);
  std::vector<long> idx;
  for (long n = 0; n < b.nnodes; n++)
    for (int d = 0; d < 3; d++)
      if (mask[3 * n + d]) {
        bits[n] |= (uint8_t)(1u << d);
        idx.push_back(3 * n + d);
      }
  PTMF_HIP(hipMemcpy(op->bcnode, bits.data(), b.nnodes, hipMemcpyHostToDevice));
  if (op->bcidx) {
    PTMF_HIP(hipFree(op->bcidx));
    op->bcidx = nullptr;
  }
  op->nbc = (long)idx.size();
  if (op->nbc) {
    PTMF_TRY(ptmf_malloc((void **)&op->bcidx, sizeof(long) * op->nbc));
    PTMF_HIP(hipMemcpy(op->bcidx, idx.data(), sizeof(long) * op->nbc, hipMemcpyHostToDevice));
  }
  op->has_bc = 1;
  op->state++;
  return PTMF_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// applies
// ---------------------------------------------------------------------------
namespace {

int check_ready(const ptmf_op_s *op, bool need_eta, bool need_bc) {
  PTMF_CHECK(op, PTMF_ERR_ARG, "null operator");
  PTMF_CHECK(op->has_coords && (!need_eta || op->has_eta) && (!need_bc || op->has_bc),
             PTMF_ERR_NOT_SETUP, "MF operator not setup");
  return PTMF_OK;
}

template <typename T>
int a11_mult(ptmf_op_s *op, const T *x, T *y) {
  PTMF_TRY(check_ready(op, true, true));
  PTMF_CHECK(x && y && (const void *)x != (const void *)y, PTMF_ERR_ARG,
             "a11_mult: x and y must be distinct device arrays");
  return ptmf::launch_a11_mult<T>(op, x, y, true);
}

}  // namespace

extern "C" {

int ptmf_a11_spmv(ptmf_op op, const double *u, double *y) {
  PTMF_TRY(check_ready(op, true, false));
  PTMF_CHECK(u && y && u != y, PTMF_ERR_ARG, "a11_spmv: u and y must be distinct device arrays");
  PTMF_TRY(ptmf::launch_a11<double>(op, u, y, false));
  PTMF_HIP(hipStreamSynchronize(op->stream));
  return PTMF_OK;
}

int ptmf_a11_spmv_host(ptmf_op op, const double *u, double *Yu) {
  PTMF_TRY(check_ready(op, true, false));
  PTMF_CHECK(u && Yu, PTMF_ERR_ARG, "a11_spmv_host: null");
  const long n3 = 3 * op->box.nnodes;
  if (!op->host_stage) PTMF_TRY(ptmf_malloc((void **)&op->host_stage, sizeof(double) * n3));
  if (!op->host_stage2) PTMF_TRY(ptmf_malloc((void **)&op->host_stage2, sizeof(double) * n3));
  // Yu is written, not uploaded and added to: the caller zeroes it right
  // before the plugin call (stokes_operators.c:1181), and the reference's
  // CUDA backend likewise overwrites it with the device result
  // (stokes_operators_cuda.cu:453), so only u crosses PCIe on the way in
  PTMF_HIP(hipMemcpyAsync(op->host_stage, u, sizeof(double) * n3, hipMemcpyHostToDevice, op->stream));
  PTMF_TRY(ptmf::launch_fill<double>(op->host_stage2, 0.0, n3, op->stream));
  PTMF_TRY(ptmf::launch_a11<double>(op, op->host_stage, op->host_stage2, false));
  PTMF_HIP(hipMemcpyAsync(Yu, op->host_stage2, sizeof(double) * n3, hipMemcpyDeviceToHost, op->stream));
  PTMF_HIP(hipStreamSynchronize(op->stream));
  return PTMF_OK;
}

// Asynchronous on the operator's stream: a Krylov loop that keeps its vectors
// on the device pays no host stall per MatMult.
int ptmf_a11_mult(ptmf_op op, const double *x, double *y) { return a11_mult<double>(op, x, y); }

int ptmf_a11_mult_f32(ptmf_op op, const float *x, float *y) { return a11_mult<float>(op, x, y); }

int ptmf_op_synchronize(ptmf_op op) {
  PTMF_CHECK(op, PTMF_ERR_ARG, "op_synchronize: null operator");
  PTMF_HIP(hipStreamSynchronize(op->stream));
  return PTMF_OK;
}

int ptmf_a11_diagonal(ptmf_op op, double *d) {
  PTMF_TRY(check_ready(op, true, true));
  PTMF_CHECK(d, PTMF_ERR_ARG, "a11_diagonal: null");
  PTMF_TRY(ptmf::launch_fill<double>(d, 0.0, 3 * op->box.nnodes, op->stream));
  PTMF_TRY(ptmf::launch_a11_diag(op, d));
  PTMF_TRY(ptmf::launch_bc_set<double>(op->bcidx, op->nbc, 1.0, d, op->stream));
  PTMF_HIP(hipStreamSynchronize(op->stream));
  return PTMF_OK;
}

int ptmf_a11_l1_diagonal(ptmf_op op, double *d) {
  PTMF_TRY(check_ready(op, true, true));
  PTMF_CHECK(d, PTMF_ERR_ARG, "a11_l1_diagonal: null");
  PTMF_TRY(ptmf::launch_fill<double>(d, 0.0, 3 * op->box.nnodes, op->stream));
  PTMF_TRY(ptmf::launch_a11_l1(op, d));
  PTMF_TRY(ptmf::launch_bc_set<double>(op->bcidx, op->nbc, 1.0, d, op->stream));
  PTMF_HIP(hipStreamSynchronize(op->stream));
  return PTMF_OK;
}

int ptmf_schur_mass_blocks(ptmf_op op, double *Me) {
  PTMF_TRY(check_ready(op, true, false));
  PTMF_CHECK(Me, PTMF_ERR_ARG, "schur_mass_blocks: null");
  PTMF_TRY(ptmf::launch_schur_blocks(op, Me, nullptr));
  PTMF_HIP(hipStreamSynchronize(op->stream));
  return PTMF_OK;
}

int ptmf_a12_mult(ptmf_op op, const double *p, double *yu) {
  PTMF_TRY(check_ready(op, false, true));
  PTMF_CHECK(p && yu, PTMF_ERR_ARG, "a12_mult: null");
  PTMF_TRY(ptmf::launch_fill<double>(yu, 0.0, 3 * op->box.nnodes, op->stream));
  PTMF_TRY(ptmf::launch_a12(op, p, yu));
  PTMF_TRY(ptmf::launch_bc_set<double>(op->bcidx, op->nbc, 0.0, yu, op->stream));
  PTMF_HIP(hipStreamSynchronize(op->stream));
  return PTMF_OK;
}

int ptmf_a21_mult(ptmf_op op, const double *u, double *yp) {
  PTMF_TRY(check_ready(op, false, true));
  PTMF_CHECK(u && yp, PTMF_ERR_ARG, "a21_mult: null");
  PTMF_TRY(ptmf::launch_a21(op, u, yp));
  PTMF_HIP(hipStreamSynchronize(op->stream));
  return PTMF_OK;
}

int ptmf_stokes_mult(ptmf_op op, const double *x, double *y) {
  PTMF_TRY(check_ready(op, true, true));
  PTMF_CHECK(x && y, PTMF_ERR_ARG, "stokes_mult: null");
  PTMF_TRY(ptmf::launch_stokes_mult(op, x, y));
  PTMF_HIP(hipStreamSynchronize(op->stream));
  return PTMF_OK;
}

}  // extern "C"

namespace {
template <typename F>
int time_loop(ptmf_op_s *op, int reps, F body, double *ms_out) {
  PTMF_HIP(hipEventRecord(op->ev0, op->stream));
  for (int r = 0; r < reps; r++) PTMF_TRY(body());
  PTMF_HIP(hipEventRecord(op->ev1, op->stream));
  PTMF_HIP(hipEventSynchronize(op->ev1));
  float ms = 0.f;
  PTMF_HIP(hipEventElapsedTime(&ms, op->ev0, op->ev1));
  if (ms_out) *ms_out = ms / reps;
  return PTMF_OK;
}
}  // namespace

extern "C" {

int ptmf_a11_time(ptmf_op op, int precision, const void *x, void *y, int reps,
                  double *ms_per_mult) {
  PTMF_CHECK(reps > 0, PTMF_ERR_ARG, "a11_time: reps must be > 0");
  PTMF_CHECK(precision == 64 || precision == 32, PTMF_ERR_UNSUPPORTED,
             "a11_time: precision %d (64 or 32)", precision);
  PTMF_TRY(check_ready(op, true, true));
  if (precision == 64)
    return time_loop(op, reps, [&] { return a11_mult<double>(op, (const double *)x, (double *)y); },
                     ms_per_mult);
  return time_loop(op, reps, [&] { return a11_mult<float>(op, (const float *)x, (float *)y); },
                   ms_per_mult);
}

int ptmf_a11_time_kernel(ptmf_op op, int precision, const void *x, void *y, int reps,
                         double *ms_per_kernel) {
  PTMF_CHECK(reps > 0, PTMF_ERR_ARG, "a11_time_kernel: reps must be > 0");
  PTMF_CHECK(precision == 64 || precision == 32, PTMF_ERR_UNSUPPORTED,
             "a11_time_kernel: precision %d (64 or 32)", precision);
  PTMF_TRY(check_ready(op, true, true));
  if (precision == 64)
    return time_loop(
        op, reps,
        [&] { return ptmf::launch_a11_mult<double>(op, (const double *)x, (double *)y, false); },
        ms_per_kernel);
  return time_loop(
      op, reps,
      [&] { return ptmf::launch_a11_mult<float>(op, (const float *)x, (float *)y, false); },
      ms_per_kernel);
}

}  // extern "C"
