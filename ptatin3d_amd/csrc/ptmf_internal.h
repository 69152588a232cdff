// ptmf_internal.h -- shared definitions of the ptmf HIP library (gfx950).
//
// Mesh model: one structured box of mx*my*mz Q2 elements, X=2mx+1, Y=2my+1,
// Z=2mz+1 nodes, node (i,j,k) -> i + X*(j + Y*k) (the DMDA numbering of
// dmda_element_q2p1.c:395-440).  Velocity vectors are node-interlaced
// (dof 3*node+d), as PETSc local vectors of the velocity DMDA.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>

#include "ptmf.h"

namespace ptmf {

// ---------------------------------------------------------------------------
// 1-D tables at the 3 Gauss-Legendre points g = (-s, 0, s), s = sqrt(3/5).
// B[q][n]: quadratic Lagrange basis (nodes -1,0,1) at g_q; D[q][n]: its
// derivative (stokes_operators_tensor.c:195-202).  C[q][r]: derivative of the
// Lagrange basis on the Gauss points themselves, L'_r(g_q) (collocation
// derivative, C = D B^{-1}); it lets the kernels interpolate once to the
// quadrature grid and differentiate there.
// ---------------------------------------------------------------------------
constexpr double kS = 0.77459666924148337703585307995647992;  // sqrt(3/5)
constexpr double kB[3][3] = {{0.3 + 0.5 * kS, 0.4, 0.3 - 0.5 * kS},
                             {0.0, 1.0, 0.0},
                             {0.3 - 0.5 * kS, 0.4, 0.3 + 0.5 * kS}};
constexpr double kD[3][3] = {{-kS - 0.5, 2.0 * kS, 0.5 - kS},
                             {-0.5, 0.0, 0.5},
                             {kS - 0.5, -2.0 * kS, kS + 0.5}};
constexpr double kC[3][3] = {{-2.5 * kS, (10.0 / 3.0) * kS, -(5.0 / 6.0) * kS},
                             {-(5.0 / 6.0) * kS, 0.0, (5.0 / 6.0) * kS},
                             {(5.0 / 6.0) * kS, -(10.0 / 3.0) * kS, 2.5 * kS}};
constexpr double kW[3] = {5.0 / 9.0, 8.0 / 9.0, 5.0 / 9.0};

struct Box {
  int mx, my, mz;
  long X, Y, Z;
  long nel, nnodes;
};

inline Box make_box(int mx, int my, int mz) {
  Box b;
  b.mx = mx; b.my = my; b.mz = mz;
  b.X = 2L * mx + 1; b.Y = 2L * my + 1; b.Z = 2L * mz + 1;
  b.nel = (long)mx * my * mz;
  b.nnodes = b.X * b.Y * b.Z;
  return b;
}

// Work decomposition of the MatMult kernel: TX x TY element tiles in x-y
// (TX*TY <= 21, one 64-lane wave each) marching through ZC layers in z.
struct Tiling {
  int TX = 0, TY = 0, ZC = 0;
  int ntx = 0, nty = 0, ntz = 0;
  long ntiles = 0;
};

// Heuristic tiling: fill the wave's 21 element slots and keep enough waves
// (about three per SIMD) in flight.
Tiling choose_tiling(int mx, int my, int mz);
Tiling make_tiling(int mx, int my, int mz, int TX, int TY, int ZC);
long face_buffer_doubles(int TX, int TY, int ZC, long ntiles);

// Element of parity class `color` (= px + 2 py + 4 pz) at position `slot` of
// the class, x fastest; no two elements of a class share a node, so their
// sums can go into a vector with plain read-add-write.  live is false for
// slots past the class (the last element is returned for them).
__host__ __device__ inline long class_element(int mx, int my, int mz, int color, long slot, bool &live) {
  const int px = color & 1, py = (color >> 1) & 1, pz = color >> 2;
  const long ncx = (mx - px + 1) / 2, ncy = (my - py + 1) / 2, ncz = (mz - pz + 1) / 2;
  live = slot < ncx * ncy * ncz;
  if (!live) slot = ncx * ncy * ncz - 1;
  return (2 * (slot % ncx) + px) + mx * ((2 * ((slot / ncx) % ncy) + py) + (long)my * (2 * (slot / (ncx * ncy)) + pz));
}

inline long class_count(const Box &b, int color) {
  return (long)((b.mx - (color & 1) + 1) / 2) * ((b.my - ((color >> 1) & 1) + 1) / 2) * ((b.mz - (color >> 2) + 1) / 2);
}

// Per-precision device copies of the operator data.
template <typename T>
struct OpData {
  T *coords = nullptr;   // 3N node-interlaced coordinates
  T *eta = nullptr;      // 27*nel, quadrature-point-major: eta[q*nel + e]
};

}  // namespace ptmf

// Operator handle (opaque in the C-ABI).
struct ptmf_op_s {
  ptmf::Box box;
  ptmf::OpData<double> d64;
  ptmf::OpData<float> d32;
  uint8_t *bcnode = nullptr;      // per-node Dirichlet bits (bit d: dof d)
  long *bcidx = nullptr;          // compact list of Dirichlet dofs
  long nbc = 0;
  int has_coords = 0, has_eta = 0, has_bc = 0;
  unsigned long state = 0;        // bumped by every coefficient change
  int device = 0;                 // device the operator was created on
  hipStream_t stream = nullptr;   // stream all work of this operator runs on
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  ptmf::Tiling tiling;
  int variant = 0;  // MatMult kernel: 0 z-march, 1 software-pipelined, 2 stored geometry, 3 coloured
  int wg_tiles = 1;  // tiles (one wave each) per work-group of the variant-0 z-march: 1, 2 or 4
  // stored quadrature-point geometry of variant 2 (per precision), valid for
  // coefficient state pa_state[i] and the tiling it was laid out for
  double *pa64 = nullptr;
  float *pa32 = nullptr;
  long pa_len = 0;
  unsigned long pa_state[2] = {~0UL, ~0UL};
  int pa_tiling[2][3] = {{0, 0, 0}, {0, 0, 0}};
  int *pa_bad = nullptr;
  // tile-ordered viscosity of variants 0/1 (per precision), same validity keys
  double *et64 = nullptr;
  float *et32 = nullptr;
  long et_len = 0;
  unsigned long et_state[2] = {~0UL, ~0UL};
  int et_tiling[2][3] = {{0, 0, 0}, {0, 0, 0}};
  void *faces = nullptr;          // tile face buffers (sized for FP64)
  long faces_doubles = 0;
  double *host_stage = nullptr;   // device staging for the host-array entry points
  double *host_stage2 = nullptr;
};

// Error plumbing: every entry point returns a ptmf status code.
int ptmf_set_error(int code, const char *fmt, ...);

#define PTMF_HIP(call)                                                              \
  do {                                                                              \
    hipError_t e_ = (call);                                                         \
    if (e_ != hipSuccess)                                                           \
      return ptmf_set_error(PTMF_ERR_DEVICE, "%s:%d %s: %s", __FILE__, __LINE__,   \
                            #call, hipGetErrorString(e_));                          \
  } while (0)

#define PTMF_CHECK(cond, code, ...)                      \
  do {                                                   \
    if (!(cond)) return ptmf_set_error(code, __VA_ARGS__); \
  } while (0)

#define PTMF_TRY(call)            \
  do {                            \
    int s_ = (call);              \
    if (s_ != PTMF_OK) return s_; \
  } while (0)

// Kernel launch helpers implemented in the .hip files (all asynchronous on
// the operator's stream).
namespace ptmf {
template <typename T>
int launch_a11(const ptmf_op_s *op, const T *u, T *y, bool mask_columns);
// MatMult_MFStokes_A11: y = A11 mask(x) on free rows, y = x on Dirichlet rows.
// faces: 1 the whole MatMult; 0 the element (tile) kernel alone, for timing;
// 2 without the Dirichlet rows (left as the element sums), for a consumer
// that takes y = x there itself.
template <typename T>
int launch_a11_mult(ptmf_op_s *op, const T *x, T *y, int faces);
// build (or rebuild) the FP64 tile-ordered viscosity for the current
// coefficients and tiling (allocates; set-up time)
int prepare_eta_tiled(ptmf_op_s *op);
template <typename T>
int launch_fill(T *x, T v, long n, hipStream_t s);
template <typename T>
int launch_bc_copy(const long *idx, long n, const T *x, T *y, hipStream_t s);
template <typename T>
int launch_bc_set(const long *idx, long n, T v, T *y, hipStream_t s);
int launch_copy16(const void *x, void *y, long n16, hipStream_t s);
int launch_eta_qpmajor(long nel, const double *em, double *q64, float *q32, hipStream_t s);
int launch_stokes_mult(ptmf_op_s *op, const double *x, double *y);
int launch_stokes_mult_f32(ptmf_op_s *op, const float *x, float *y);
// element blocks (-B_e^T D^-1 B_e)^-1; *singular (device, zeroed by the
// caller) is set when a block has fewer than 4 independent free rows
int launch_schur_bdb(const ptmf_op_s *op, const double *dg, double *Minv, int *singular);
// batched MatMult on device tables of per-box x / y pointers (precision 32 or
// 64): batch_bind builds the work list and box descriptors once (and the
// tile-ordered viscosity), batch_apply launches on stream s (capturable);
// boxes whose active[box] is 0 are skipped (active may be null: all boxes)
int batch_bind(ptmf_batch_s *b, int precision);
int batch_apply(ptmf_batch_s *b, int precision, const void *const *xs, void *const *ys, hipStream_t s,
                const int *active = nullptr);
int launch_powerlaw(const ptmf_op_s *op, const double *u, const double *B, double nexp, double eta_min,
                    double eta_max, double *eta);
int launch_a11_diag(const ptmf_op_s *op, double *d);
int launch_a11_l1(const ptmf_op_s *op, double *d);
int launch_schur_blocks(const ptmf_op_s *op, double *Me, double *Minv);
int launch_schur_apply(const ptmf_op_s *op, const double *Minv, const double *rp, double *yp);
int launch_a12(const ptmf_op_s *op, const double *p, double *yu);
// Multifrontal Cholesky of a level operator given as its 125-point 3x3-block
// stencil S (device, S[((o*9)+3r+c)*N + m] = A(3m+r, 3(m+off_o)+c), SPD with
// unit Dirichlet rows), nested-dissection ordered (coarse.hip).  mf_solve is
// kernels only (capturable, no allocation); b and x are device FP64 arrays.
struct MFFactor;
// bc_host: per-node Dirichlet bits on the host (nodes with all three set are
// left out of the factor) or null
int mf_factor(const Box &box, const double *S, const uint8_t *bc_host, hipStream_t stream, MFFactor **out);
int mf_solve(const MFFactor *m, const double *b, double *x, hipStream_t s);
void mf_destroy(MFFactor *m);
long mf_factor_doubles(const MFFactor *m);
double mf_setup_ms(const MFFactor *m);
int mf_depths(const MFFactor *m);
int launch_a21(const ptmf_op_s *op, const double *u, double *yp);
}  // namespace ptmf
