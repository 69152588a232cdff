// a11.hip -- matrix-free Q2 viscous block (A11) apply for gfx950.
//
// Computes A11_local u for all elements of a structured Q2 box, the operator
// of MFStokesWrapper_A11_Tensor (stokes_operators_tensor.c:181-288): per
// element and quadrature point, grad u and the Jacobian from the 27 nodal
// values, then w*detJ*(2 eta sym(grad u)) G^T, then the transposed gradient
// back to the nodes.
//
// Element work decomposition (MI355X-first, not one-warp-per-element):
//   * 3 lanes per element, lane c owns quadrature plane zeta_c (9 qps) and,
//     on input/output, node plane k = c of the element; 21 elements per
//     64-lane wave, one wave per work-group (no cross-wave barriers).
//   * Collocated sum factorisation: nodal -> quadrature-grid interpolation
//     (B x B x B), then the three reference derivatives on the quadrature
//     grid with the 3x3 collocation matrix C = D B^-1.  The xi/eta stages
//     are in registers; only the zeta stage crosses lanes, through LDS.
//   * The 1/detJ of the inverse Jacobian is folded into one scalar per qp
//     (adjugate form): one division per qp.
//   * Viscosity is stored quadrature-point-major (eta[q*nel + e]) so the
//     lanes of a wave read consecutive elements.
//
// Scatter (MatMult path, a11_tile_kernel + a11_face_kernel):
//   * A wave owns a TX x TY tile of elements in x-y (TX*TY <= 21) and marches
//     through a chunk of ZC element layers in z.  Element outputs are summed
//     into an LDS window of three node planes (ordered plain read-add-write,
//     LDS float atomics are slow here); after each layer the two completed
//     node planes are flushed with plain stores.
//   * A tile owns its node range [2x0,2x1) x [2y0,2y1) x [2z0,2z1) (plus the
//     closing planes at the domain's upper boundary).  Contributions to the
//     tile's upper faces go to per-tile face buffers (x face first, then y,
//     then z), and a11_face_kernel adds them into the owning tile's lower
//     faces.  No global atomics, no clearing of y, deterministic sums.
//   * Dirichlet rows (y = x) are written last, over the compact list of
//     Dirichlet dofs.
// The spmv path (y += A u, no boundary handling) keeps a one-pass kernel with
// global atomic adds; it serves the host-array PETSc adapter contract.
#include <string.h>

#include <vector>

#include "ptmf_internal.h"

namespace ptmf {
namespace {

constexpr int kEPW = 21;  // elements per wave (3 lanes each; lane 63 idle)
constexpr int kSlots = 54;
// Exchange-buffer columns: the 21 element slots, plus (FP32) one for the
// idle lane 63, which then reads and writes a column of its own and needs
// no write guards.  Measured on the 64^3 MatMult: the FP32 kernel is 3-4 %
// faster with the spare column, the FP64 kernel 2.5 % slower, so FP64 keeps
// 21 columns and guards its writes.
template <typename T>
constexpr int sh_cols() { return sizeof(T) == 4 ? kEPW + 1 : kEPW; }
constexpr int kWinNodes = 130;  // >= (2TX+1)(2TY+1) for every TX*TY <= 21

// Barrier of one element tile's lanes.  A work-group holds kW waves, each
// marching through its own tile with its own LDS buffers, so with kW > 1 the
// waves never wait for each other (their tiles may have different numbers
// of layers): the barrier is the wave's own (LDS operations of one wave
// execute in order; the fences keep the compiler from moving them across).
template <int kWaves>
__device__ __forceinline__ void tsync() {
  if constexpr (kWaves == 1) {
    __syncthreads();
  } else {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
}

// lane of the calling thread within its wave
template <int kWaves>
__device__ __forceinline__ int wave_lane() {
  if constexpr (kWaves == 1) return (int)threadIdx.x;
  else return (int)(threadIdx.x & 63);
}

template <typename T>
__device__ __forceinline__ T sel3(int c, double a0, double a1, double a2) {
  return (T)(c == 0 ? a0 : (c == 1 ? a1 : a2));
}

// Raw input of one lane's node plane: coordinates and velocity of the 9
// nodes (i, j) of plane k = cp, their Dirichlet bits, and the viscosity at
// the lane's 9 quadrature points.  Kept unmodified in registers between the
// load and the forward stage so the loads can be issued a layer ahead.
template <typename T>
struct PlaneRaw {
  T x[3][9];  // [d][3j+i]
  T u[3][9];
  unsigned m[9];
  T eta[9];
};

template <typename T, bool kMask>
__device__ __forceinline__ void load_plane_nodes(PlaneRaw<T> &r, const T *__restrict__ xc,
                                                 const T *__restrict__ u,
                                                 const uint8_t *__restrict__ bcn, long base, long X) {
#pragma unroll
  for (int j = 0; j < 3; j++)
#pragma unroll
    for (int i = 0; i < 3; i++) {
      const long node = base + i + X * j;
      r.m[3 * j + i] = kMask ? (unsigned)bcn[node] : 0u;
#pragma unroll
      for (int d = 0; d < 3; d++) {
        r.x[d][3 * j + i] = xc[3 * node + d];
        r.u[d][3 * j + i] = u[3 * node + d];
      }
    }
}

template <typename T>
__device__ __forceinline__ void load_plane_eta(PlaneRaw<T> &r, const T *__restrict__ etap, long stride) {
#pragma unroll
  for (int p = 0; p < 9; p++) r.eta[p] = etap[(long)p * stride];
}

// 1-D operators written with the exact structure of the Gauss tables (the
// compiler may not drop multiplications by exact zeros or ones under IEEE
// rules): B = [[e+o, m, e-o], [0, 1, 0], [e-o, m, e+o]] with e = 0.3,
// o = sqrt(3/5)/2, m = 0.4, evaluated in even/odd form (6 operations instead
// of 9).  interp3: y = B x; interp3t: y = B^T x.
constexpr double kBe = 0.5 * (kB[0][0] + kB[0][2]);
constexpr double kBo = 0.5 * (kB[0][0] - kB[0][2]);

template <typename T>
__device__ __forceinline__ void interp3(T x0, T x1, T x2, T &y0, T &y1, T &y2) {
  const T sm = x0 + x2, df = x0 - x2;
  const T t = (T)kB[0][1] * x1 + (T)kBe * sm;
  y0 = t + (T)kBo * df;
  y1 = x1;
  y2 = t - (T)kBo * df;
}

template <typename T>
__device__ __forceinline__ void interp3t(T x0, T x1, T x2, T &y0, T &y1, T &y2) {
  const T sm = x0 + x2, df = x0 - x2;
  const T t = (T)kBe * sm;
  y0 = t + (T)kBo * df;
  y1 = x1 + (T)kB[0][1] * sm;
  y2 = t - (T)kBo * df;
}

// collocated derivative at quadrature point a of a 3-point line,
// sum_r C[a][r] x_r, with the zero of the middle row skipped
template <typename T>
__device__ __forceinline__ T cdot(int a, T x0, T x1, T x2) {
  if (a == 1) return (T)kC[1][2] * (x2 - x0);
  return (T)kC[a][0] * x0 + (T)kC[a][1] * x1 + (T)kC[a][2] * x2;
}

// 1/x by the hardware reciprocal estimate refined with Newton-Raphson steps
// r <- r + r (1 - x r) (two for double, one for float)
template <typename T>
__device__ __forceinline__ T recip(T x);
template <>
__device__ __forceinline__ double recip<double>(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = fma(-x, r, 1.0);
  r = fma(r, e, r);
  e = fma(-x, r, 1.0);
  return fma(r, e, r);
}
template <>
__device__ __forceinline__ float recip<float>(float x) {
  const float r = __builtin_amdgcn_rcpf(x);
  const float e = fmaf(-x, r, 1.0f);
  return fmaf(r, e, r);
}

// in-plane interpolation of a 3x3 node plane to the 3x3 quadrature grid:
// V[d][3b+a] = sum_{i,j} B[a][i] B[b][j] g[d][3j+i]
template <typename T>
__device__ __forceinline__ void interp_plane(const T g[3][9], T V[3][9]) {
#pragma unroll
  for (int d = 0; d < 3; d++) {
    T t[3][3];  // [j][a]
#pragma unroll
    for (int j = 0; j < 3; j++) interp3<T>(g[d][3 * j], g[d][3 * j + 1], g[d][3 * j + 2], t[j][0], t[j][1], t[j][2]);
#pragma unroll
    for (int a = 0; a < 3; a++) interp3<T>(t[0][a], t[1][a], t[2][a], V[d][a], V[d][3 + a], V[d][6 + a]);
  }
}

// transposed in-plane interpolation: out[3j+i] = sum_{a,b} B[a][i] B[b][j] Y[3b+a]
template <typename T>
__device__ __forceinline__ void interp_plane_t(const T Y[9], T out[9]) {
  T t[3][3];  // [b][i]
#pragma unroll
  for (int b = 0; b < 3; b++) interp3t<T>(Y[3 * b], Y[3 * b + 1], Y[3 * b + 2], t[b][0], t[b][1], t[b][2]);
#pragma unroll
  for (int i = 0; i < 3; i++) interp3t<T>(t[0][i], t[1][i], t[2][i], out[i], out[3 + i], out[6 + i]);
}

// Forward stage: in-plane interpolation of the lane's node plane, exchange
// through the LDS buffer (slots 0..26 coordinates, 27..53 masked velocity),
// zeta interpolation to the lane's quadrature plane.
template <typename T, bool kMask, int kWaves = 1>
__device__ __forceinline__ void plane_forward(int cp, int le, bool writer, const PlaneRaw<T> &r,
                                              T (*sh)[kSlots][sh_cols<T>()], T Xq[3][9], T Uq[3][9]) {
  T bz[3];
#pragma unroll
  for (int k = 0; k < 3; k++) bz[k] = sel3<T>(cp, kB[0][k], kB[1][k], kB[2][k]);
  {
    T V[3][9];
    interp_plane<T>(r.x, V);
    if (writer) {
#pragma unroll
      for (int d = 0; d < 3; d++)
#pragma unroll
        for (int p = 0; p < 9; p++) sh[cp][9 * d + p][le] = V[d][p];
    }
  }
  {
    T g[3][9], V[3][9];
#pragma unroll
    for (int d = 0; d < 3; d++)
#pragma unroll
      for (int n = 0; n < 9; n++) g[d][n] = ((r.m[n] >> d) & 1u) ? T(0) : r.u[d][n];
    interp_plane<T>(g, V);
    if (writer) {
#pragma unroll
      for (int d = 0; d < 3; d++)
#pragma unroll
        for (int p = 0; p < 9; p++) sh[cp][27 + 9 * d + p][le] = V[d][p];
    }
  }
  tsync<kWaves>();
#pragma unroll
  for (int d = 0; d < 3; d++)
#pragma unroll
    for (int p = 0; p < 9; p++) {
      Xq[d][p] = bz[0] * sh[0][9 * d + p][le] + bz[1] * sh[1][9 * d + p][le] +
                 bz[2] * sh[2][9 * d + p][le];
      Uq[d][p] = bz[0] * sh[0][27 + 9 * d + p][le] + bz[1] * sh[1][27 + 9 * d + p][le] +
                 bz[2] * sh[2][27 + 9 * d + p][le];
    }
}

// Pointwise stage on the 9 quadrature points of the lane's plane: Jacobian
// (adjugate form, one division per point), symmetric gradient, viscous
// stress, and the transposed derivatives.  The xi/eta parts are accumulated
// in W and exchanged with the zeta parts through the LDS buffer: on return,
// slots 9l+p hold the zeta components and slots 27+9l+p the xi/eta
// accumulations of every plane.
//
// A work-group is a single wave, so the LDS operations of all lanes execute
// in program order: a slot read by every lane at qp p is overwritten only
// after those reads.
// Per-lane data of the pressure coupling in the fused saddle-point MatMult:
// the element centroid xb, the pressure coefficients with the element
// lengths divided out, p_h(x) = pc0 + sum_d pc[1+d] (x_d - xb_d) (the P1GRel
// basis, element_utils_q2.c:183-224), and the lane's partial sums of
// B^T u against {1, x - xb, y - yb, z - zb}.
template <typename T>
struct PCoupling {
  T xb[3];
  T pc[4];
  T acc[4];
};

template <typename T, bool kStokes = false, int kWaves = 1>
__device__ __forceinline__ void plane_qp(int cp, int le, bool writer, const T Xq[3][9],
                                         const T Uq[3][9], const T etq9[9],
                                         T (*sh)[kSlots][sh_cols<T>()], PCoupling<T> *pcp = nullptr) {
  T dz[3];
#pragma unroll
  for (int k = 0; k < 3; k++) dz[k] = sel3<T>(cp, kD[0][k], kD[1][k], kD[2][k]);
  const T wz = sel3<T>(cp, kW[0], kW[1], kW[2]);
  T W[3][9];  // in-plane transposed xi/eta derivatives, accumulated
#pragma unroll
  for (int l = 0; l < 3; l++)
#pragma unroll
    for (int p = 0; p < 9; p++) W[l][p] = T(0);

  // zeta-plane values of the next point are read before this point's zeta
  // part is written (point p writes slots p, 9+p, 18+p; point p+1 reads
  // slots p+1, 9+p+1, ... only)
  T Z[2][2][3][3];
#pragma unroll
  for (int d = 0; d < 3; d++)
#pragma unroll
    for (int k = 0; k < 3; k++) {
      Z[0][0][d][k] = sh[k][9 * d][le];
      Z[0][1][d][k] = sh[k][27 + 9 * d][le];
    }
#pragma unroll
  for (int b = 0; b < 3; b++)
#pragma unroll
    for (int a = 0; a < 3; a++) {
      const int p = 3 * b + a;
      if (p < 8) {
#pragma unroll
        for (int d = 0; d < 3; d++)
#pragma unroll
          for (int k = 0; k < 3; k++) {
            Z[(p + 1) & 1][0][d][k] = sh[k][9 * d + p + 1][le];
            Z[(p + 1) & 1][1][d][k] = sh[k][27 + 9 * d + p + 1][le];
          }
      }
      const T etq = etq9[p];
      // M[j][d] = d x_d / d xi_j, Gu[j][l] = d u_l / d xi_j
      T M[3][3], Gu[3][3];
#pragma unroll
      for (int d = 0; d < 3; d++) {
        M[0][d] = cdot<T>(a, Xq[d][3 * b], Xq[d][3 * b + 1], Xq[d][3 * b + 2]);
        M[1][d] = cdot<T>(b, Xq[d][a], Xq[d][3 + a], Xq[d][6 + a]);
        M[2][d] = dz[0] * Z[p & 1][0][d][0] + dz[1] * Z[p & 1][0][d][1] + dz[2] * Z[p & 1][0][d][2];
        Gu[0][d] = cdot<T>(a, Uq[d][3 * b], Uq[d][3 * b + 1], Uq[d][3 * b + 2]);
        Gu[1][d] = cdot<T>(b, Uq[d][a], Uq[d][3 + a], Uq[d][6 + a]);
        Gu[2][d] = dz[0] * Z[p & 1][1][d][0] + dz[1] * Z[p & 1][1][d][1] + dz[2] * Z[p & 1][1][d][2];
      }
      // adjugate A = det * M^-1 (A[k][j] = det * d xi_j / d x_k)
      T A[3][3];
      A[0][0] = M[1][1] * M[2][2] - M[2][1] * M[1][2];
      A[0][1] = M[2][1] * M[0][2] - M[0][1] * M[2][2];
      A[0][2] = M[0][1] * M[1][2] - M[1][1] * M[0][2];
      A[1][0] = M[2][0] * M[1][2] - M[1][0] * M[2][2];
      A[1][1] = M[0][0] * M[2][2] - M[2][0] * M[0][2];
      A[1][2] = M[1][0] * M[0][2] - M[0][0] * M[1][2];
      A[2][0] = M[1][0] * M[2][1] - M[2][0] * M[1][1];
      A[2][1] = M[2][0] * M[0][1] - M[0][0] * M[2][1];
      A[2][2] = M[0][0] * M[1][1] - M[1][0] * M[0][1];
      const T det = M[0][0] * A[0][0] + M[0][1] * A[1][0] + M[0][2] * A[2][0];
      const T wq = (T)(kW[a] * kW[b]) * wz;
      const T s = (T(2) * wq * etq) * recip<T>(det);
      // gh[k][l] = det * d u_l / d x_k
      T gh[3][3];
#pragma unroll
      for (int k = 0; k < 3; k++)
#pragma unroll
        for (int l = 0; l < 3; l++)
          gh[k][l] = Gu[0][l] * A[k][0] + Gu[1][l] * A[k][1] + Gu[2][l] * A[k][2];
      // sigma = 2 eta sym(grad u) * w * detJ / det^2 (scaled so that the
      // adjugate can stand in for detJ * G)
      const T hs = T(0.5) * s;
      T sg[3][3];
      sg[0][0] = s * gh[0][0];
      sg[1][1] = s * gh[1][1];
      sg[2][2] = s * gh[2][2];
      sg[0][1] = sg[1][0] = hs * (gh[0][1] + gh[1][0]);
      sg[0][2] = sg[2][0] = hs * (gh[0][2] + gh[2][0]);
      sg[1][2] = sg[2][1] = hs * (gh[1][2] + gh[2][1]);
      if constexpr (kStokes) {
        // A12: the pressure enters the stress as -p_h I; A21: the lane's
        // share of -sum_q w detJ NIp_i div u (w detJ div u = w tr(gh))
        const T d0 = Xq[0][p] - pcp->xb[0], d1 = Xq[1][p] - pcp->xb[1], d2 = Xq[2][p] - pcp->xb[2];
        const T wp = wq * (pcp->pc[0] + pcp->pc[1] * d0 + pcp->pc[2] * d1 + pcp->pc[3] * d2);
        sg[0][0] -= wp;
        sg[1][1] -= wp;
        sg[2][2] -= wp;
        const T dv = -wq * (gh[0][0] + gh[1][1] + gh[2][2]);
        pcp->acc[0] += dv;
        pcp->acc[1] += dv * d0;
        pcp->acc[2] += dv * d1;
        pcp->acc[3] += dv * d2;
      }
      // every lane has read slot p of all planes; reuse it for the zeta data
      __builtin_amdgcn_wave_barrier();
      // dv[j][l] = w detJ sum_m sigma[m][l] d xi_j / d x_m
#pragma unroll
      for (int l = 0; l < 3; l++) {
        const T dv0 = sg[0][l] * A[0][0] + sg[1][l] * A[1][0] + sg[2][l] * A[2][0];
        const T dv1 = sg[0][l] * A[0][1] + sg[1][l] * A[1][1] + sg[2][l] * A[2][1];
        const T dv2 = sg[0][l] * A[0][2] + sg[1][l] * A[1][2] + sg[2][l] * A[2][2];
#pragma unroll
        for (int r = 0; r < 3; r++) {
          if (a != 1 || r != 1) W[l][3 * b + r] += (T)kC[a][r] * dv0;
          if (b != 1 || r != 1) W[l][3 * r + a] += (T)kC[b][r] * dv1;
        }
        if (writer) sh[cp][9 * l + p][le] = dv2;  // zeta component, exchanged below
      }
      __builtin_amdgcn_wave_barrier();
    }
  tsync<kWaves>();
  if (writer) {
#pragma unroll
    for (int l = 0; l < 3; l++)
#pragma unroll
      for (int p = 0; p < 9; p++) sh[cp][27 + 9 * l + p][le] = W[l][p];
  }
  tsync<kWaves>();
}

// Backward stage: node plane k = cp gathers the three quadrature planes
// (zeta), then the transposed in-plane interpolation; out[l][3j+i] is the
// element result at node (i, j, k = cp), component l.  Ends with a barrier:
// the exchange buffer is free again on return.
template <typename T, int kWaves = 1>
__device__ __forceinline__ void plane_backward(int cp, int le, T (*sh)[kSlots][sh_cols<T>()], T out[3][9]) {
  T bt[3], dt[3];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    bt[k] = sel3<T>(cp, kB[k][0], kB[k][1], kB[k][2]);
    dt[k] = sel3<T>(cp, kD[k][0], kD[k][1], kD[k][2]);
  }
  T Y[3][9];
#pragma unroll
  for (int l = 0; l < 3; l++)
#pragma unroll
    for (int p = 0; p < 9; p++) {
      const T z0 = sh[0][9 * l + p][le], z1 = sh[1][9 * l + p][le], z2 = sh[2][9 * l + p][le];
      const T w0 = sh[0][27 + 9 * l + p][le], w1 = sh[1][27 + 9 * l + p][le],
              w2 = sh[2][27 + 9 * l + p][le];
      Y[l][p] = bt[0] * w0 + bt[1] * w1 + bt[2] * w2 + dt[0] * z0 + dt[1] * z1 + dt[2] * z2;
    }
  tsync<kWaves>();
#pragma unroll
  for (int l = 0; l < 3; l++) interp_plane_t<T>(Y[l], out[l]);
}

// The element computation of one lane: element e (node plane origin `base`,
// lane plane cp), result out[l][3j+i] = (A11_e u)(node (i,j,k=cp), comp l).
// Every lane of the wave must call it (it contains wave-wide barriers); lanes
// with writer == false (lane 63) never write the exchange buffer.
template <typename T, bool kMask, int kWaves = 1>
__device__ __forceinline__ void element_plane_apply(int cp, int le, bool writer, const T *__restrict__ etap,
                                                    long estride, long base, long X,
                                                    const T *__restrict__ xc,
                                                    const uint8_t *__restrict__ bcn,
                                                    const T *__restrict__ u,
                                                    T (*sh)[kSlots][sh_cols<T>()], T out[3][9]) {
  PlaneRaw<T> r;
  load_plane_nodes<T, kMask>(r, xc, u, bcn, base, X);
  load_plane_eta<T>(r, etap, estride);
  T Xq[3][9], Uq[3][9];
  plane_forward<T, kMask, kWaves>(cp, le, writer, r, sh, Xq, Uq);
  plane_qp<T, false, kWaves>(cp, le, writer, Xq, Uq, r.eta, sh);
  plane_backward<T, kWaves>(cp, le, sh, out);
}

// The same element computation for the saddle-point operator
// [A11 A12; A21 0] (MatMult_MFStokes_A, stokes_operators.c:884-967): the
// element's pressure pe[0..3] enters the stress, and B^T of the masked
// velocity is written to ype[0..3] by the lane of plane 0 (live lanes only).
// sL holds the element lengths between the first and last stage.
template <typename T>
__device__ __forceinline__ void element_plane_apply_stokes(int cp, int le, bool writer, bool live,
                                                           const T *__restrict__ etap, long estride, long base,
                                                           long X, const T *__restrict__ xc,
                                                           const uint8_t *__restrict__ bcn,
                                                           const T *__restrict__ u, const T *__restrict__ pe,
                                                           T *__restrict__ ype, T (*sh)[kSlots][sh_cols<T>()],
                                                           T (*sL)[kEPW], T out[3][9]) {
  PlaneRaw<T> r;
  load_plane_nodes<T, true>(r, xc, u, bcn, base, X);
  load_plane_eta<T>(r, etap, estride);
  PCoupling<T> pcp;
  // centroid and lengths: x14 - x12 and y16 - y10 lie in plane 1, z22 - z4
  // joins the centre nodes of planes 2 and 0
  if (writer) {
#pragma unroll
    for (int d = 0; d < 3; d++) {
      T sum = T(0);
#pragma unroll
      for (int n = 0; n < 9; n++) sum += r.x[d][n];
      sh[cp][d][le] = sum;
    }
    sh[cp][3][le] = r.x[2][4];
    sh[cp][4][le] = r.x[0][5] - r.x[0][3];
    sh[cp][5][le] = r.x[1][7] - r.x[1][1];
  }
  __syncthreads();
  const T Lx = sh[1][4][le], Ly = sh[1][5][le], Lz = sh[2][3][le] - sh[0][3][le];
#pragma unroll
  for (int d = 0; d < 3; d++) pcp.xb[d] = (sh[0][d][le] + sh[1][d][le] + sh[2][d][le]) * T(1.0 / 27.0);
  if (writer && cp == 0) {
    sL[0][le] = Lx;
    sL[1][le] = Ly;
    sL[2][le] = Lz;
  }
  pcp.pc[0] = pe[0];
  pcp.pc[1] = pe[1] / Lx;
  pcp.pc[2] = pe[2] / Ly;
  pcp.pc[3] = pe[3] / Lz;
#pragma unroll
  for (int i = 0; i < 4; i++) pcp.acc[i] = T(0);
  __syncthreads();
  T Xq[3][9], Uq[3][9];
  plane_forward<T, true>(cp, le, writer, r, sh, Xq, Uq);
  plane_qp<T, true>(cp, le, writer, Xq, Uq, r.eta, sh, &pcp);
  plane_backward<T>(cp, le, sh, out);
  if (writer) {
#pragma unroll
    for (int i = 0; i < 4; i++) sh[cp][i][le] = pcp.acc[i];
  }
  __syncthreads();
  if (live && cp == 0) {
    ype[0] = sh[0][0][le] + sh[1][0][le] + sh[2][0][le];
#pragma unroll
    for (int i = 1; i < 4; i++) ype[i] = (sh[0][i][le] + sh[1][i][le] + sh[2][i][le]) / sL[i - 1][le];
  }
  __syncthreads();
}

// ---------------------------------------------------------------------------
// spmv: y += A11_local u with global atomic adds (no boundary handling)
// ---------------------------------------------------------------------------
template <typename T, bool kMask>
__global__ void __launch_bounds__(64)
    a11_atomic_kernel(int mx, int my, long nel, long X, long XY, const T *__restrict__ xc,
                      const T *__restrict__ eta, const uint8_t *__restrict__ bcn,
                      const T *__restrict__ u, T *__restrict__ y) {
  __shared__ T sh[3][kSlots][sh_cols<T>()];
  const int lane = threadIdx.x;
  const int c = lane / kEPW;  // quadrature / node plane; 3 marks the idle lane
  const int le = lane - c * kEPW;
  const int cp = c < 3 ? c : 2;
  long e = (long)blockIdx.x * kEPW + le;
  const bool live = (c < 3) && (e < nel);
  if (e >= nel) e = nel - 1;
  const long ei = e % mx, er = e / mx, ej = er % my, ek = er / my;
  const long base = 2 * ei + X * (2 * ej) + XY * (2 * ek + cp);
  T out[3][9];
  element_plane_apply<T, kMask>(cp, le, c < 3, eta + (long)(9 * cp) * nel + e, nel, base, X, xc, bcn,
                                u, sh, out);
  if (live) {
#pragma unroll
    for (int l = 0; l < 3; l++)
#pragma unroll
      for (int j = 0; j < 3; j++)
#pragma unroll
        for (int i = 0; i < 3; i++) atomicAdd(&y[3 * (base + i + X * j) + l], out[l][3 * j + i]);
  }
}

// ---------------------------------------------------------------------------
// MatMult with a coloured scatter (kernel variant 3): one launch per element
// parity class, 21 elements of the class per wave, each lane adding its
// node plane's sums into y with plain read-add-write (no two elements of a
// class share a node).  y is cleared first; viscosity read quadrature-point
// major (eta[q * nel + e]).
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(64)
    a11_colour_kernel(int mx, int my, int mz, long nel, long X, long XY, int color, const T *__restrict__ xc,
                      const T *__restrict__ eta, const uint8_t *__restrict__ bcn, const T *__restrict__ u,
                      T *__restrict__ y) {
  __shared__ T sh[3][kSlots][sh_cols<T>()];
  const int lane = threadIdx.x;
  const int c = lane / kEPW;  // quadrature / node plane; 3 marks the idle lane
  const int le = lane - c * kEPW;
  const int cp = c < 3 ? c : 2;
  bool in_class;
  const long e = class_element(mx, my, mz, color, (long)blockIdx.x * kEPW + le, in_class);
  const bool live = (c < 3) && in_class;
  const long ei = e % mx, er = e / mx, ej = er % my, ek = er / my;
  const long base = 2 * ei + X * (2 * ej) + XY * (2 * ek + cp);
  T out[3][9];
  element_plane_apply<T, true>(cp, le, c < 3, eta + (long)(9 * cp) * nel + e, nel, base, X, xc, bcn, u, sh, out);
  if (live) {
#pragma unroll
    for (int j = 0; j < 3; j++)
#pragma unroll
      for (int i = 0; i < 3; i++)
#pragma unroll
        for (int l = 0; l < 3; l++) {
          T *q = y + 3 * (base + i + X * j) + l;
          *q += out[l][3 * j + i];
        }
  }
}

// ---------------------------------------------------------------------------
// MatMult: tiled z-march with an LDS node-plane window and face buffers
// ---------------------------------------------------------------------------
struct TileArgs {
  int mx, my, mz;
  long X, XY, nel;
  int TX, TY, ZC, ntx, nty;
  long ntiles;
  long fbx_n, fby_n, fbz_n;  // per-tile face buffer sizes in nodes
};

// Work-groups are dispatched round-robin over the 8 XCDs (each with its own
// L2); give every XCD a contiguous run of tiles (x fastest), so tiles that
// share halo nodes and viscosity cache lines share an L2.
__device__ __forceinline__ long xcd_tile(long b, long ntiles) {
  const long per = ntiles / 8, rem = ntiles % 8;
  const long x = b % 8, slot = b / 8;
  return x * per + (x < rem ? x : rem) + slot;
}

struct TileBox {
  int tx, ty, tz, x0, y0, z0, nxt, nyt, nzt;
  bool xend, yend, zend;
};

__device__ __forceinline__ TileBox tile_box(const TileArgs &a, long tile) {
  TileBox t;
  t.tx = (int)(tile % a.ntx);
  t.ty = (int)((tile / a.ntx) % a.nty);
  t.tz = (int)(tile / ((long)a.ntx * a.nty));
  t.x0 = t.tx * a.TX;
  t.y0 = t.ty * a.TY;
  t.z0 = t.tz * a.ZC;
  t.nxt = min(a.TX, a.mx - t.x0);
  t.nyt = min(a.TY, a.my - t.y0);
  t.nzt = min(a.ZC, a.mz - t.z0);
  t.xend = t.x0 + t.nxt == a.mx;
  t.yend = t.y0 + t.nyt == a.my;
  t.zend = t.z0 + t.nzt == a.mz;
  return t;
}

// Tile-ordered quadrature-point data (viscosity for variants 0/1): for tile
// t, layer kz of its z-chunk, quadrature point q and element slot le of the
// wave, index ((t * ZC + kz) * 27 + q) * 21 + le, so the lanes of a plane read
// 21 consecutive values.  Slots beyond a partial tile repeat the clamped
// element, as the MatMult kernels do.
__device__ __forceinline__ long tq_index(long tile, int ZC, int kz, int q, int le) {
  return ((tile * ZC + kz) * 27 + q) * (long)kEPW + le;
}

template <typename T>
__global__ void __launch_bounds__(256)
    eta_tile_kernel(TileArgs a, const double *__restrict__ eta, T *__restrict__ et) {
  const long total = a.ntiles * a.ZC * 27L * kEPW;
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < total; t += (long)gridDim.x * blockDim.x) {
    const int le = (int)(t % kEPW);
    const int q = (int)((t / kEPW) % 27);
    const int kz = (int)((t / (27L * kEPW)) % a.ZC);
    const long tile = t / (27L * kEPW * a.ZC);
    const TileBox tb = tile_box(a, tile);
    if (kz >= tb.nzt) continue;
    const int lx = le % a.TX, ly = le / a.TX;
    const int ex = tb.x0 + min(lx, tb.nxt - 1), ey = tb.y0 + min(ly, tb.nyt - 1);
    const long e = ex + (long)a.mx * (ey + (long)a.my * (tb.z0 + kz));
    et[t] = (T)eta[(long)q * a.nel + e];
  }
}

// Write the completed node plane P of the window: owned nodes to y, the
// tile's upper-face nodes to the face buffers (the x face column, corner
// included, then the y face row, then on the tile's top plane the z face).
// Pure LDS reads and global stores; clears the window entries it reads.
// Three loops, one destination each (no divergent stores): the owned block
// (node columns il < nxo, rows jl < nyo), the upper x face column and the
// upper y face row.
template <typename T, int kWaves = 1>
__device__ __forceinline__ void flush_plane(const TileArgs &a, long tile, const TileBox &tb, int P,
                                            T *__restrict__ win, T *__restrict__ y,
                                            T *__restrict__ fbx, T *__restrict__ fby,
                                            T *__restrict__ fbz) {
  const int wX = 2 * tb.nxt + 1, wY = 2 * tb.nyt + 1;
  const int W3 = 3 * wX;
  const int kl = P - 2 * tb.z0;
  const int nxo = tb.xend ? wX : 2 * tb.nxt;  // owned node columns
  const int nyo = tb.yend ? wY : 2 * tb.nyt;  // owned node rows
  const int R3 = 3 * nxo;
  const int lane = wave_lane<kWaves>();
  {
    // owned block: entry t = jl * R3 + r (r = 3 il + l) is window entry jl * W3 + r
    const bool zface = kl >= 2 * tb.nzt && !tb.zend;
    T *const dst = zface ? fbz : y;
    const long off = zface ? 3 * tile * a.fbz_n
                           : 3 * ((long)(2 * tb.x0) + a.XY * (long)P + a.X * (long)(2 * tb.y0));
    const long rs = zface ? 3L * (2 * a.TX + 1) : 3 * a.X;
    const float inv = 1.0f / (float)R3;
    for (int t = lane; t < R3 * nyo; t += 64) {
      int jl = (int)((float)t * inv);
      if (jl * R3 > t) jl--;
      if ((jl + 1) * R3 <= t) jl++;
      const int r = t - jl * R3;
      const int w = jl * W3 + r;
      const T v = win[w];
      win[w] = T(0);
      dst[off + rs * jl + r] = v;
    }
  }
  if (!tb.xend) {
    // upper x face: node column il = 2 nxt of every row, t = 3 jl + l
    const long off = 3 * (tile * a.fbx_n + (long)(2 * a.TY + 1) * kl);
    for (int t = lane; t < 3 * wY; t += 64) {
      const int jl = t / 3;
      const int w = jl * W3 + 6 * tb.nxt + (t - 3 * jl);
      const T v = win[w];
      win[w] = T(0);
      fbx[off + t] = v;
    }
  }
  if (!tb.yend) {
    // upper y face: node row jl = 2 nyt without the x face corner, t = 3 il + l
    const long off = 3 * (tile * a.fby_n + (long)(2 * a.TX + 1) * kl);
    for (int t = lane; t < R3; t += 64) {
      const int w = 2 * tb.nyt * W3 + t;
      const T v = win[w];
      win[w] = T(0);
      fby[off + t] = v;
    }
  }
}

// One tile's z-march (the work of one wave); shared by the single-box and
// the batched MatMult kernels.
template <typename T, bool kMask, bool kStokes = false, int kWaves = 1>
__device__ __forceinline__ void tile_march(const TileArgs &a, long tile, const T *__restrict__ xc,
                                           const T *__restrict__ eta, const uint8_t *__restrict__ bcn,
                                           const T *__restrict__ u, T *__restrict__ y, T *__restrict__ fbx,
                                           T *__restrict__ fby, T *__restrict__ fbz,
                                           T (*sh)[kSlots][sh_cols<T>()], T (*win)[3 * kWinNodes],
                                           const T *__restrict__ p = nullptr, T *__restrict__ yp = nullptr,
                                           T (*sL)[kEPW] = nullptr) {
  const TileBox tb = tile_box(a, tile);
  const int lane = wave_lane<kWaves>();
  const int c = lane / kEPW;
  const int le = lane - c * kEPW;
  const int cp = c < 3 ? c : 2;
  const int lx = le % a.TX, ly = le / a.TX;
  const bool live = (c < 3) && (le < a.TX * a.TY) && (lx < tb.nxt) && (ly < tb.nyt);
  const int ex = tb.x0 + min(lx, tb.nxt - 1), ey = tb.y0 + min(ly, tb.nyt - 1);
  const int wX = 2 * tb.nxt + 1;
  // exchange-buffer column and write permission: with the spare column
  // (FP32) the idle lane computes on element slot 0 but uses a column of its
  // own, so every lane may write
  constexpr bool kSpare = sh_cols<T>() > kEPW;
  const int lc = (c < 3 || !kSpare) ? le : kEPW;
  const bool wr = kSpare || c < 3;

  for (int t = lane; t < 9 * kWinNodes; t += 64) (&win[0][0])[t] = T(0);
  tsync<kWaves>();

  for (int kz = 0; kz < tb.nzt; kz++) {
    const long ek = tb.z0 + kz;
    const long base = 2L * ex + a.X * (2L * ey) + a.XY * (2 * ek + cp);
    T out[3][9];
    if constexpr (kStokes) {
      const long e = ex + (long)a.mx * (ey + (long)a.my * ek);
      element_plane_apply_stokes<T>(cp, le, c < 3, live, eta + tq_index(tile, a.ZC, kz, 9 * cp, le), kEPW, base,
                                    a.X, xc, bcn, u, p + 4 * e, yp + 4 * e, sh, sL, out);
    } else {
      element_plane_apply<T, kMask, kWaves>(cp, lc, wr, eta + tq_index(tile, a.ZC, kz, 9 * cp, le), kEPW, base,
                                            a.X, xc, bcn, u, sh, out);
    }
    // Sum into the window with plain read-add-write.  Element (lx, ly)
    // touches nodes (2lx + i, 2ly + j); only slots i = 2 / j = 2 meet the
    // slots i = 0 / j = 0 of the next element, so the nine slots go in four
    // groups, within which no two lanes touch the same entry: {i, j < 2},
    // {i = 2, j < 2}, {i < 2, j = 2}, {(2, 2)}.  The wave executes LDS
    // operations in order; the barriers between groups keep the compiler
    // from interleaving them.
    {
      T *w = win[(2 * kz + cp) % 3] + 3 * ((2 * ly) * wX + 2 * lx);
#pragma unroll
      for (int g = 0; g < 4; g++) {
#pragma unroll
        for (int j = 0; j < 3; j++)
#pragma unroll
          for (int i = 0; i < 3; i++) {
            if ((i == 2 ? 1 : 0) + (j == 2 ? 2 : 0) != g) continue;
            if (live) {
              T *q = w + 3 * (j * wX + i);
              const T q0 = q[0], q1 = q[1], q2 = q[2];
              q[0] = q0 + out[0][3 * j + i];
              q[1] = q1 + out[1][3 * j + i];
              q[2] = q2 + out[2][3 * j + i];
            }
          }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
      }
    }
    tsync<kWaves>();
    flush_plane<T, kWaves>(a, tile, tb, (int)(2 * ek), win[(2 * kz) % 3], y, fbx, fby, fbz);
    flush_plane<T, kWaves>(a, tile, tb, (int)(2 * ek + 1), win[(2 * kz + 1) % 3], y, fbx, fby, fbz);
    tsync<kWaves>();
  }
  flush_plane<T, kWaves>(a, tile, tb, 2 * (tb.z0 + tb.nzt), win[(2 * tb.nzt) % 3], y, fbx, fby, fbz);
}

template <typename T, bool kMask>
__global__ void __launch_bounds__(64)
    a11_tile_kernel(TileArgs a, const T *__restrict__ xc, const T *__restrict__ eta,
                    const uint8_t *__restrict__ bcn, const T *__restrict__ u, T *__restrict__ y,
                    T *__restrict__ fbx, T *__restrict__ fby, T *__restrict__ fbz) {
  __shared__ T sh[3][kSlots][sh_cols<T>()];
  __shared__ T win[3][3 * kWinNodes];
  tile_march<T, kMask>(a, xcd_tile(blockIdx.x, a.ntiles), xc, eta, bcn, u, y, fbx, fby, fbz, sh, win);
}

// The same z-march with kW tiles per work-group (one wave each, own LDS
// buffers): work-group g takes tiles kW g .. kW g + kW - 1 of the x-fastest
// order, work-groups mapped to XCDs as above.  A wave past the last tile
// returns at once; no barrier spans waves.
template <typename T, bool kMask, int kWaves>
__global__ void __launch_bounds__(64 * kWaves) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 2 : 1)))
    a11_tile_wg_kernel(TileArgs a, long nwg, const T *__restrict__ xc, const T *__restrict__ eta,
                       const uint8_t *__restrict__ bcn, const T *__restrict__ u, T *__restrict__ y,
                       T *__restrict__ fbx, T *__restrict__ fby, T *__restrict__ fbz) {
  __shared__ T sh[kWaves][3][kSlots][sh_cols<T>()];
  __shared__ T win[kWaves][3][3 * kWinNodes];
  const int w = (int)(threadIdx.x >> 6);
  const long tile = xcd_tile(blockIdx.x, nwg) * kWaves + w;
  if (tile >= a.ntiles) return;
  tile_march<T, kMask, false, kWaves>(a, tile, xc, eta, bcn, u, y, fbx, fby, fbz, sh[w], win[w]);
}

// Saddle-point MatMult: the A11 z-march with the pressure coupling; y holds
// the velocity part, yp the pressure part (4 per element, written once).
template <typename T>
__global__ void __launch_bounds__(64)
    stokes_tile_kernel(TileArgs a, const T *__restrict__ xc, const T *__restrict__ eta,
                       const uint8_t *__restrict__ bcn, const T *__restrict__ u, const T *__restrict__ p,
                       T *__restrict__ y, T *__restrict__ yp, T *__restrict__ fbx, T *__restrict__ fby,
                       T *__restrict__ fbz) {
  __shared__ T sh[3][kSlots][sh_cols<T>()];
  __shared__ T win[3][3 * kWinNodes];
  __shared__ T sL[3][kEPW];
  tile_march<T, true, true>(a, xcd_tile(blockIdx.x, a.ntiles), xc, eta, bcn, u, y, fbx, fby, fbz, sh, win, p, yp,
                            sL);
}

// The same z-march, software-pipelined across layers: the raw node plane of
// layer k+1 is loaded while layer k's backward stage runs, its viscosity
// while layer k's window sums run, and layer k's completed node planes are
// flushed after layer k+1's forward stage, so no global load is waited on
// right after it is issued.
template <typename T, bool kMask>
__global__ void __launch_bounds__(64)
    a11_tile_pf_kernel(TileArgs a, const T *__restrict__ xc, const T *__restrict__ eta,
                       const uint8_t *__restrict__ bcn, const T *__restrict__ u, T *__restrict__ y,
                       T *__restrict__ fbx, T *__restrict__ fby, T *__restrict__ fbz) {
  __shared__ T sh[3][kSlots][sh_cols<T>()];
  __shared__ T win[3][3 * kWinNodes];
  const long tile = xcd_tile(blockIdx.x, a.ntiles);
  const TileBox tb = tile_box(a, tile);
  const int lane = threadIdx.x;
  const int c = lane / kEPW;
  const int le = lane - c * kEPW;
  const int cp = c < 3 ? c : 2;
  const bool writer = c < 3;
  const int lx = le % a.TX, ly = le / a.TX;
  const bool live = writer && (le < a.TX * a.TY) && (lx < tb.nxt) && (ly < tb.nyt);
  const int ex = tb.x0 + min(lx, tb.nxt - 1), ey = tb.y0 + min(ly, tb.nyt - 1);
  const int wX = 2 * tb.nxt + 1;
  const long ncol = 2L * ex + a.X * (2L * ey);  // node (0, 0, 0) of element (ex, ey, 0)

  PlaneRaw<T> r;
  load_plane_nodes<T, kMask>(r, xc, u, bcn, ncol + a.XY * (2L * tb.z0 + cp), a.X);
  load_plane_eta<T>(r, eta + tq_index(tile, a.ZC, 0, 9 * cp, le), kEPW);
  for (int t = lane; t < 9 * kWinNodes; t += 64) (&win[0][0])[t] = T(0);
  __syncthreads();

  for (int kz = 0; kz < tb.nzt; kz++) {
    const long ek = tb.z0 + kz;
    T out[3][9];
    {
      T Xq[3][9], Uq[3][9];
      plane_forward<T, kMask>(cp, le, writer, r, sh, Xq, Uq);
      plane_qp<T>(cp, le, writer, Xq, Uq, r.eta, sh);
    }
    plane_backward<T>(cp, le, sh, out);
    // next layer's loads: in flight during the window sums and the flush
    __builtin_amdgcn_sched_barrier(0);
    if (kz + 1 < tb.nzt) {
      load_plane_nodes<T, kMask>(r, xc, u, bcn, ncol + a.XY * (2 * ek + 2 + cp), a.X);
      load_plane_eta<T>(r, eta + tq_index(tile, a.ZC, kz + 1, 9 * cp, le), kEPW);
    }
    __builtin_amdgcn_sched_barrier(0);
    // window sums, as in a11_tile_kernel
    {
      T *w = win[(2 * kz + cp) % 3] + 3 * ((2 * ly) * wX + 2 * lx);
#pragma unroll
      for (int j = 0; j < 3; j++)
#pragma unroll
        for (int i = 0; i < 3; i++) {
          if (live) {
            T *q = w + 3 * (j * wX + i);
            const T q0 = q[0], q1 = q[1], q2 = q[2];
            q[0] = q0 + out[0][3 * j + i];
            q[1] = q1 + out[1][3 * j + i];
            q[2] = q2 + out[2][3 * j + i];
          }
          __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
          __builtin_amdgcn_wave_barrier();
        }
    }
    __syncthreads();
    flush_plane<T>(a, tile, tb, (int)(2 * ek), win[(2 * kz) % 3], y, fbx, fby, fbz);
    flush_plane<T>(a, tile, tb, (int)(2 * ek + 1), win[(2 * kz + 1) % 3], y, fbx, fby, fbz);
    __syncthreads();
  }
  flush_plane<T>(a, tile, tb, 2 * (tb.z0 + tb.nzt), win[(2 * tb.nzt) % 3], y, fbx, fby, fbz);
}

// ---------------------------------------------------------------------------
// Partial assembly (kernel variant 2): per quadrature point the 9 entries of
// Ah = sqrt(2 w eta / det) * adj(J) (adj(J) = det * J^-1, rows k: d xi / d x_k
// scaled), stored once per coefficient/tiling change.  The pointwise action
// of MFStokesWrapper_A11_Tensor (QuadratureAction, tensor.c:127-179) is then
//   dv = sym(Gu Ah^T) Ah,   Gu[j][l] = d u_l / d xi_j,
// which equals w detJ (2 eta sym(grad u)) G^T with G = J^-1.  Layout follows
// the MatMult tiles so a lane group reads 21 consecutive values:
//   pa[((((tile * ZC + kz) * 27 + q) * 9 + m) * 21) + le].
// ---------------------------------------------------------------------------
__device__ __forceinline__ long pa_index(long tile, int ZC, int kz, int q, int m, int le) {
  return ((((tile * ZC + kz) * 27 + q) * 9 + m) * (long)kEPW) + le;
}

template <typename T>
__global__ void __launch_bounds__(256)
    pa_build_kernel(TileArgs a, long ntiles, const double *__restrict__ xc,
                    const double *__restrict__ eta, T *__restrict__ pa, int *__restrict__ bad) {
  const long total = ntiles * a.ZC * 27L * kEPW;
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < total; t += (long)gridDim.x * blockDim.x) {
    const int le = (int)(t % kEPW);
    const int q = (int)((t / kEPW) % 27);
    const int kz = (int)((t / (27L * kEPW)) % a.ZC);
    const long tile = t / (27L * kEPW * a.ZC);
    const TileBox tb = tile_box(a, tile);
    if (kz >= tb.nzt) continue;
    const int lx = le % a.TX, ly = le / a.TX;
    const int ex = tb.x0 + min(lx, tb.nxt - 1), ey = tb.y0 + min(ly, tb.nyt - 1);
    const long ek = tb.z0 + kz;
    const long e = ex + (long)a.mx * (ey + (long)a.my * ek);
    const int qa = q % 3, qb = (q / 3) % 3, qc = q / 9;
    // M[j][d] = d x_d / d xi_j at the quadrature point
    double M[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    const long base = 2L * ex + a.X * (2L * ey) + a.XY * (2L * ek);
    for (int k = 0; k < 3; k++)
      for (int j = 0; j < 3; j++)
        for (int i = 0; i < 3; i++) {
          const long n = base + i + a.X * j + a.XY * k;
          const double g0 = kD[qa][i] * kB[qb][j] * kB[qc][k];
          const double g1 = kB[qa][i] * kD[qb][j] * kB[qc][k];
          const double g2 = kB[qa][i] * kB[qb][j] * kD[qc][k];
          for (int d = 0; d < 3; d++) {
            const double x = xc[3 * n + d];
            M[0][d] += g0 * x;
            M[1][d] += g1 * x;
            M[2][d] += g2 * x;
          }
        }
    double A[3][3];
    A[0][0] = M[1][1] * M[2][2] - M[2][1] * M[1][2];
    A[0][1] = M[2][1] * M[0][2] - M[0][1] * M[2][2];
    A[0][2] = M[0][1] * M[1][2] - M[1][1] * M[0][2];
    A[1][0] = M[2][0] * M[1][2] - M[1][0] * M[2][2];
    A[1][1] = M[0][0] * M[2][2] - M[2][0] * M[0][2];
    A[1][2] = M[1][0] * M[0][2] - M[0][0] * M[1][2];
    A[2][0] = M[1][0] * M[2][1] - M[2][0] * M[1][1];
    A[2][1] = M[2][0] * M[0][1] - M[0][0] * M[2][1];
    A[2][2] = M[0][0] * M[1][1] - M[1][0] * M[0][1];
    const double det = M[0][0] * A[0][0] + M[0][1] * A[1][0] + M[0][2] * A[2][0];
    const double s = 2.0 * kW[qa] * kW[qb] * kW[qc] * eta[(long)q * a.nel + e] / det;
    if (!(s > 0.0) || !isfinite(s)) *bad = 1;
    const double r = sqrt(fabs(s));
    for (int k = 0; k < 3; k++)
      for (int j = 0; j < 3; j++) pa[pa_index(tile, a.ZC, kz, q, 3 * k + j, le)] = (T)(r * A[k][j]);
  }
}

// Pointwise stage with stored geometry (see above); same exchange protocol
// as plane_qp with the velocity in slots 0..26 of a 27-slot buffer: on
// return slots 9l+p hold the zeta components, W the xi/eta accumulations.
template <typename T>
__device__ __forceinline__ void plane_qp_pa(int cp, int le, bool writer, const T Uq[3][9],
                                            const T *__restrict__ pal, T (*sh)[27][kEPW],
                                            T W[3][9]) {
  T dz[3];
#pragma unroll
  for (int k = 0; k < 3; k++) dz[k] = sel3<T>(cp, kD[0][k], kD[1][k], kD[2][k]);
#pragma unroll
  for (int l = 0; l < 3; l++)
#pragma unroll
    for (int p = 0; p < 9; p++) W[l][p] = T(0);
  T An[9];
#pragma unroll
  for (int m = 0; m < 9; m++) An[m] = pal[m * kEPW];
#pragma unroll
  for (int b = 0; b < 3; b++)
#pragma unroll
    for (int a = 0; a < 3; a++) {
      const int p = 3 * b + a;
      T Ah[9];
#pragma unroll
      for (int m = 0; m < 9; m++) Ah[m] = An[m];
      if (p < 8) {
#pragma unroll
        for (int m = 0; m < 9; m++) An[m] = pal[((p + 1) * 9 + m) * kEPW];
      }
      T Gu[3][3];
#pragma unroll
      for (int d = 0; d < 3; d++) {
        Gu[0][d] = cdot<T>(a, Uq[d][3 * b], Uq[d][3 * b + 1], Uq[d][3 * b + 2]);
        Gu[1][d] = cdot<T>(b, Uq[d][a], Uq[d][3 + a], Uq[d][6 + a]);
        Gu[2][d] = dz[0] * sh[0][9 * d + p][le] + dz[1] * sh[1][9 * d + p][le] +
                   dz[2] * sh[2][9 * d + p][le];
      }
      // gh[k][l] = sum_j Gu[j][l] Ah[k][j]; sg = sym(gh)
      T gh[3][3];
#pragma unroll
      for (int k = 0; k < 3; k++)
#pragma unroll
        for (int l = 0; l < 3; l++)
          gh[k][l] = Gu[0][l] * Ah[3 * k] + Gu[1][l] * Ah[3 * k + 1] + Gu[2][l] * Ah[3 * k + 2];
      T sg[3][3];
      sg[0][0] = gh[0][0];
      sg[1][1] = gh[1][1];
      sg[2][2] = gh[2][2];
      sg[0][1] = sg[1][0] = T(0.5) * (gh[0][1] + gh[1][0]);
      sg[0][2] = sg[2][0] = T(0.5) * (gh[0][2] + gh[2][0]);
      sg[1][2] = sg[2][1] = T(0.5) * (gh[1][2] + gh[2][1]);
      __builtin_amdgcn_wave_barrier();
#pragma unroll
      for (int l = 0; l < 3; l++) {
        const T dv0 = sg[0][l] * Ah[0] + sg[1][l] * Ah[3] + sg[2][l] * Ah[6];
        const T dv1 = sg[0][l] * Ah[1] + sg[1][l] * Ah[4] + sg[2][l] * Ah[7];
        const T dv2 = sg[0][l] * Ah[2] + sg[1][l] * Ah[5] + sg[2][l] * Ah[8];
#pragma unroll
        for (int r = 0; r < 3; r++) {
          if (a != 1 || r != 1) W[l][3 * b + r] += (T)kC[a][r] * dv0;
          if (b != 1 || r != 1) W[l][3 * r + a] += (T)kC[b][r] * dv1;
        }
        if (writer) sh[cp][9 * l + p][le] = dv2;
      }
      __builtin_amdgcn_wave_barrier();
    }
  __syncthreads();
}

// MatMult with stored geometry: the z-march of a11_tile_kernel with only the
// velocity exchanged (27 LDS slots) and the backward zeta stage in two
// passes over the same slots.
template <typename T, bool kMask>
__global__ void __launch_bounds__(64)
    a11_tile_pa_kernel(TileArgs a, const T *__restrict__ pa, const uint8_t *__restrict__ bcn,
                       const T *__restrict__ u, T *__restrict__ y, T *__restrict__ fbx,
                       T *__restrict__ fby, T *__restrict__ fbz) {
  __shared__ T sh[3][27][kEPW];
  __shared__ T win[3][3 * kWinNodes];
  const long tile = xcd_tile(blockIdx.x, a.ntiles);
  const TileBox tb = tile_box(a, tile);
  const int lane = threadIdx.x;
  const int c = lane / kEPW;
  const int le = lane - c * kEPW;
  const int cp = c < 3 ? c : 2;
  const bool writer = c < 3;
  const int lx = le % a.TX, ly = le / a.TX;
  const bool live = writer && (le < a.TX * a.TY) && (lx < tb.nxt) && (ly < tb.nyt);
  const int ex = tb.x0 + min(lx, tb.nxt - 1), ey = tb.y0 + min(ly, tb.nyt - 1);
  const int wX = 2 * tb.nxt + 1;
  const long ncol = 2L * ex + a.X * (2L * ey);

  for (int t = lane; t < 9 * kWinNodes; t += 64) (&win[0][0])[t] = T(0);
  __syncthreads();

  T bz[3], bt[3], dt[3];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    bz[k] = sel3<T>(cp, kB[0][k], kB[1][k], kB[2][k]);
    bt[k] = sel3<T>(cp, kB[k][0], kB[k][1], kB[k][2]);
    dt[k] = sel3<T>(cp, kD[k][0], kD[k][1], kD[k][2]);
  }

  for (int kz = 0; kz < tb.nzt; kz++) {
    const long ek = tb.z0 + kz;
    const long base = ncol + a.XY * (2 * ek + cp);
    const T *pal = pa + pa_index(tile, a.ZC, kz, 9 * cp, 0, le);
    // forward: masked velocity plane, in-plane interpolation, zeta exchange
    T Uq[3][9];
    {
      T g[3][9];
#pragma unroll
      for (int j = 0; j < 3; j++)
#pragma unroll
        for (int i = 0; i < 3; i++) {
          const long node = base + i + a.X * j;
          const unsigned m = kMask ? (unsigned)bcn[node] : 0u;
#pragma unroll
          for (int d = 0; d < 3; d++) {
            const T v = u[3 * node + d];
            g[d][3 * j + i] = ((m >> d) & 1u) ? T(0) : v;
          }
        }
      T V[3][9];
      interp_plane<T>(g, V);
      if (writer) {
#pragma unroll
        for (int d = 0; d < 3; d++)
#pragma unroll
          for (int p = 0; p < 9; p++) sh[cp][9 * d + p][le] = V[d][p];
      }
    }
    __syncthreads();
#pragma unroll
    for (int d = 0; d < 3; d++)
#pragma unroll
      for (int p = 0; p < 9; p++)
        Uq[d][p] = bz[0] * sh[0][9 * d + p][le] + bz[1] * sh[1][9 * d + p][le] +
                   bz[2] * sh[2][9 * d + p][le];
    T W[3][9];
    plane_qp_pa<T>(cp, le, writer, Uq, pal, sh, W);
    // backward zeta stage, pass 1: the zeta components
    T Y[3][9];
#pragma unroll
    for (int l = 0; l < 3; l++)
#pragma unroll
      for (int p = 0; p < 9; p++)
        Y[l][p] = dt[0] * sh[0][9 * l + p][le] + dt[1] * sh[1][9 * l + p][le] +
                  dt[2] * sh[2][9 * l + p][le];
    __syncthreads();
    if (writer) {
#pragma unroll
      for (int l = 0; l < 3; l++)
#pragma unroll
        for (int p = 0; p < 9; p++) sh[cp][9 * l + p][le] = W[l][p];
    }
    __syncthreads();
    // pass 2: the xi/eta accumulations
#pragma unroll
    for (int l = 0; l < 3; l++)
#pragma unroll
      for (int p = 0; p < 9; p++)
        Y[l][p] += bt[0] * sh[0][9 * l + p][le] + bt[1] * sh[1][9 * l + p][le] +
                   bt[2] * sh[2][9 * l + p][le];
    __syncthreads();
    T out[3][9];
#pragma unroll
    for (int l = 0; l < 3; l++) interp_plane_t<T>(Y[l], out[l]);
    {
      T *w = win[(2 * kz + cp) % 3] + 3 * ((2 * ly) * wX + 2 * lx);
#pragma unroll
      for (int j = 0; j < 3; j++)
#pragma unroll
        for (int i = 0; i < 3; i++) {
          if (live) {
            T *q = w + 3 * (j * wX + i);
            const T q0 = q[0], q1 = q[1], q2 = q[2];
            q[0] = q0 + out[0][3 * j + i];
            q[1] = q1 + out[1][3 * j + i];
            q[2] = q2 + out[2][3 * j + i];
          }
          __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
          __builtin_amdgcn_wave_barrier();
        }
    }
    __syncthreads();
    flush_plane<T>(a, tile, tb, (int)(2 * ek), win[(2 * kz) % 3], y, fbx, fby, fbz);
    flush_plane<T>(a, tile, tb, (int)(2 * ek + 1), win[(2 * kz + 1) % 3], y, fbx, fby, fbz);
    __syncthreads();
  }
  flush_plane<T>(a, tile, tb, 2 * (tb.z0 + tb.nzt), win[(2 * tb.nzt) % 3], y, fbx, fby, fbz);
}

// Adds the neighbours' face-buffer contributions into the nodes on the lower
// faces of each tile (the work-group's threads take one node each in turn;
// every node's sum has one thread and a fixed order).
template <typename T>
__device__ __forceinline__ void face_sum(const TileArgs &a, long tile, const T *__restrict__ fbx,
                                         const T *__restrict__ fby, const T *__restrict__ fbz,
                                         T *__restrict__ y) {
  const TileBox tb = tile_box(a, tile);
  const int oX = 2 * tb.nxt + (tb.xend ? 1 : 0);
  const int oY = 2 * tb.nyt + (tb.yend ? 1 : 0);
  const int oZ = 2 * tb.nzt + (tb.zend ? 1 : 0);
  const bool hx = tb.tx > 0, hy = tb.ty > 0, hz = tb.tz > 0;
  const int nX = hx ? oY * oZ : 0;
  const int iy0 = hx ? 1 : 0, jz0 = hy ? 1 : 0, iz0 = hx ? 1 : 0;
  const int nY = hy ? (oX - iy0) * oZ : 0;
  const int nZ = hz ? (oX - iz0) * (oY - jz0) : 0;
  for (int t = threadIdx.x; t < nX + nY + nZ; t += blockDim.x) {
    int il, jl, kl;
    if (t < nX) {
      il = 0; jl = t % oY; kl = t / oY;
    } else if (t < nX + nY) {
      const int r = t - nX, w = oX - iy0;
      il = iy0 + r % w; jl = 0; kl = r / w;
    } else {
      const int r = t - nX - nY, w = oX - iz0;
      il = iz0 + r % w; jl = jz0 + r / w; kl = 0;
    }
    const int I = 2 * tb.x0 + il, J = 2 * tb.y0 + jl, K = 2 * tb.z0 + kl;
    T acc[3] = {T(0), T(0), T(0)};
    for (int dz = (kl == 0 && hz) ? -1 : 0; dz <= 0; dz++)
      for (int dy = (jl == 0 && hy) ? -1 : 0; dy <= 0; dy++)
        for (int dx = (il == 0 && hx) ? -1 : 0; dx <= 0; dx++) {
          if (dx == 0 && dy == 0 && dz == 0) continue;
          const long tp = (tb.tx + dx) + (long)a.ntx * ((tb.ty + dy) + (long)a.nty * (tb.tz + dz));
          const int ilp = I - 2 * (tb.tx + dx) * a.TX, jlp = J - 2 * (tb.ty + dy) * a.TY,
                    klp = K - 2 * (tb.tz + dz) * a.ZC;
          const T *src;
          if (dx == -1)
            src = fbx + 3 * (tp * a.fbx_n + jlp + (long)(2 * a.TY + 1) * klp);
          else if (dy == -1)
            src = fby + 3 * (tp * a.fby_n + ilp + (long)(2 * a.TX + 1) * klp);
          else
            src = fbz + 3 * (tp * a.fbz_n + ilp + (long)(2 * a.TX + 1) * jlp);
          acc[0] += src[0];
          acc[1] += src[1];
          acc[2] += src[2];
        }
    const long n = (long)I + a.X * (long)J + a.XY * (long)K;
#pragma unroll
    for (int l = 0; l < 3; l++) y[3 * n + l] += acc[l];
  }
}

// One tile per work-group of kFaceThreads, one face node per thread: the
// tall tiles of long z-chunks have 400-800 lower-face nodes, which a single
// wave would walk one dependent load/store round after another.
constexpr int kFaceThreads = 256;

template <typename T>
__global__ void __launch_bounds__(kFaceThreads)
    a11_face_kernel(TileArgs a, const T *__restrict__ fbx, const T *__restrict__ fby,
                    const T *__restrict__ fbz, T *__restrict__ y) {
  face_sum<T>(a, xcd_tile(blockIdx.x, a.ntiles), fbx, fby, fbz, y);
}

// ---------------------------------------------------------------------------
// Batched MatMult: many independent boxes in one launch per stage.  A work
// item is (box, tile); every box has its own tiling, coefficients, Dirichlet
// data and face buffers (BoxDesc, in device memory).
// ---------------------------------------------------------------------------
struct BoxDesc {
  TileArgs a;
  const void *xc, *eta;  // coordinates and tile-ordered viscosity of the precision
  const uint8_t *bcn;
  const long *bcidx;
  long nbc;
  void *fbx, *fby, *fbz;
};

template <typename T>
__global__ void __launch_bounds__(64)
    a11_tile_batch_kernel(const BoxDesc *__restrict__ boxes, const int2 *__restrict__ work, long nwork,
                          const T *const *__restrict__ xs, T *const *__restrict__ ys,
                          const int *__restrict__ active) {
  __shared__ T sh[3][kSlots][sh_cols<T>()];
  __shared__ T win[3][3 * kWinNodes];
  const int2 w = work[xcd_tile(blockIdx.x, nwork)];
  if (active && !active[w.x]) return;  // whole work-group: the box is idle
  const BoxDesc &d = boxes[w.x];
  tile_march<T, true>(d.a, w.y, (const T *)d.xc, (const T *)d.eta, d.bcn, xs[w.x], ys[w.x], (T *)d.fbx,
                      (T *)d.fby, (T *)d.fbz, sh, win);
}

template <typename T>
__global__ void __launch_bounds__(64)
    a11_face_batch_kernel(const BoxDesc *__restrict__ boxes, const int2 *__restrict__ work, long nwork,
                          T *const *__restrict__ ys, const int *__restrict__ active) {
  const int2 w = work[xcd_tile(blockIdx.x, nwork)];
  if (active && !active[w.x]) return;
  const BoxDesc &d = boxes[w.x];
  face_sum<T>(d.a, w.y, (const T *)d.fbx, (const T *)d.fby, (const T *)d.fbz, ys[w.x]);
}

// Dirichlet rows y = x of every box (blockIdx.y = box)
template <typename T>
__global__ void bc_copy_batch_kernel(const BoxDesc *__restrict__ boxes, const T *const *__restrict__ xs,
                                     T *const *__restrict__ ys, const int *__restrict__ active) {
  if (active && !active[blockIdx.y]) return;
  const BoxDesc &d = boxes[blockIdx.y];
  const T *x = xs[blockIdx.y];
  T *y = ys[blockIdx.y];
  for (long k = blockIdx.x * (long)blockDim.x + threadIdx.x; k < d.nbc; k += (long)gridDim.x * blockDim.x) {
    const long g = d.bcidx[k];
    y[g] = x[g];
  }
}

TileArgs tile_args(const ptmf_op_s *op) {
  const Box &b = op->box;
  TileArgs a;
  a.mx = b.mx; a.my = b.my; a.mz = b.mz;
  a.X = b.X; a.XY = b.X * b.Y; a.nel = b.nel;
  a.TX = op->tiling.TX; a.TY = op->tiling.TY; a.ZC = op->tiling.ZC;
  a.ntx = op->tiling.ntx; a.nty = op->tiling.nty;
  a.ntiles = op->tiling.ntiles;
  a.fbx_n = (long)(2 * a.TY + 1) * (2 * a.ZC + 1);
  a.fby_n = (long)(2 * a.TX + 1) * (2 * a.ZC + 1);
  a.fbz_n = (long)(2 * a.TX + 1) * (2 * a.TY + 1);
  return a;
}

}  // namespace

long face_buffer_doubles(int TX, int TY, int ZC, long ntiles) {
  const long per = (long)(2 * TY + 1) * (2 * ZC + 1) + (long)(2 * TX + 1) * (2 * ZC + 1) +
                   (long)(2 * TX + 1) * (2 * TY + 1);
  return 3 * per * ntiles;
}

template <typename T>
int launch_a11(const ptmf_op_s *op, const T *u, T *y, bool mask_columns) {
  const Box &b = op->box;
  const OpData<T> &d = *(const OpData<T> *)(sizeof(T) == 8 ? (const void *)&op->d64
                                                           : (const void *)&op->d32);
  const long nblk = (b.nel + kEPW - 1) / kEPW;
  PTMF_CHECK(nblk < (1L << 31), PTMF_ERR_ARG, "a11: mesh too large (%ld blocks)", nblk);
  if (mask_columns)
    a11_atomic_kernel<T, true><<<dim3((unsigned)nblk), dim3(64), 0, op->stream>>>(
        b.mx, b.my, b.nel, b.X, b.X * b.Y, d.coords, d.eta, op->bcnode, u, y);
  else
    a11_atomic_kernel<T, false><<<dim3((unsigned)nblk), dim3(64), 0, op->stream>>>(
        b.mx, b.my, b.nel, b.X, b.X * b.Y, d.coords, d.eta, nullptr, u, y);
  PTMF_HIP(hipGetLastError());
  return PTMF_OK;
}

// (Re)build the tile-ordered viscosity for precision T when the coefficients
// or the tiling changed since the last build.
template <typename T>
int ensure_eta_tiled(ptmf_op_s *op) {
  const int ip = sizeof(T) == 8 ? 0 : 1;
  const Tiling &tl = op->tiling;
  if (op->et_state[ip] == op->state && op->et_tiling[ip][0] == tl.TX &&
      op->et_tiling[ip][1] == tl.TY && op->et_tiling[ip][2] == tl.ZC)
    return PTMF_OK;
  const long len = tl.ntiles * tl.ZC * 27L * kEPW;
  if (len > op->et_len) {
    PTMF_HIP(hipStreamSynchronize(op->stream));
    (void)hipFree(op->et64);
    (void)hipFree(op->et32);
    op->et64 = nullptr;
    op->et32 = nullptr;
    op->et_state[0] = op->et_state[1] = ~0UL;
    op->et_len = len;
  }
  T **dst = (T **)(sizeof(T) == 8 ? (void *)&op->et64 : (void *)&op->et32);
  if (!*dst) PTMF_HIP(hipMalloc((void **)dst, sizeof(T) * op->et_len));
  const TileArgs a = tile_args(op);
  long nb = (len + 255) / 256;
  if (nb > 65536) nb = 65536;
  eta_tile_kernel<T><<<dim3((unsigned)nb), dim3(256), 0, op->stream>>>(a, op->d64.eta, *dst);
  PTMF_HIP(hipGetLastError());
  op->et_state[ip] = op->state;
  op->et_tiling[ip][0] = tl.TX;
  op->et_tiling[ip][1] = tl.TY;
  op->et_tiling[ip][2] = tl.ZC;
  return PTMF_OK;
}

// (Re)build the stored geometry of variant 2 for precision T when the
// coefficients or the tiling changed since the last build.
template <typename T>
int ensure_pa(ptmf_op_s *op) {
  const int ip = sizeof(T) == 8 ? 0 : 1;
  const Tiling &tl = op->tiling;
  if (op->pa_state[ip] == op->state && op->pa_tiling[ip][0] == tl.TX &&
      op->pa_tiling[ip][1] == tl.TY && op->pa_tiling[ip][2] == tl.ZC)
    return PTMF_OK;
  const long len = tl.ntiles * tl.ZC * 27L * 9L * kEPW;
  if (len > op->pa_len) {
    PTMF_HIP(hipStreamSynchronize(op->stream));
    (void)hipFree(op->pa64);
    (void)hipFree(op->pa32);
    op->pa64 = nullptr;
    op->pa32 = nullptr;
    op->pa_state[0] = op->pa_state[1] = ~0UL;
    op->pa_len = len;
  }
  T **dst = (T **)(sizeof(T) == 8 ? (void *)&op->pa64 : (void *)&op->pa32);
  if (!*dst) PTMF_HIP(hipMalloc((void **)dst, sizeof(T) * op->pa_len));
  if (!op->pa_bad) PTMF_HIP(hipMalloc((void **)&op->pa_bad, sizeof(int)));
  PTMF_TRY(launch_fill<int>(op->pa_bad, 0, 1, op->stream));
  const TileArgs a = tile_args(op);
  const long total = tl.ntiles * tl.ZC * 27L * kEPW;
  long nb = (total + 255) / 256;
  if (nb > 65536) nb = 65536;
  pa_build_kernel<T><<<dim3((unsigned)nb), dim3(256), 0, op->stream>>>(a, tl.ntiles, op->d64.coords,
                                                                      op->d64.eta, *dst, op->pa_bad);
  PTMF_HIP(hipGetLastError());
  int bad = 0;
  PTMF_HIP(hipMemcpyAsync(&bad, op->pa_bad, sizeof(int), hipMemcpyDeviceToHost, op->stream));
  PTMF_HIP(hipStreamSynchronize(op->stream));
  PTMF_CHECK(!bad, PTMF_ERR_ARG,
             "a11 (stored geometry): non-positive or non-finite 2 w eta / det at some quadrature "
             "point (inverted element or eta <= 0)");
  op->pa_state[ip] = op->state;
  op->pa_tiling[ip][0] = tl.TX;
  op->pa_tiling[ip][1] = tl.TY;
  op->pa_tiling[ip][2] = tl.ZC;
  return PTMF_OK;
}

// The tile-ordered viscosity for precision T rebuilt, when stale, into the
// buffer the operator already holds (kernel only, no allocation); an error
// when that buffer does not exist or is too short for the current tiling.
template <typename T>
int eta_tiled_in_place(ptmf_op_s *op) {
  const int ip = sizeof(T) == 8 ? 0 : 1;
  const Tiling &tl = op->tiling;
  if (op->et_state[ip] == op->state && op->et_tiling[ip][0] == tl.TX && op->et_tiling[ip][1] == tl.TY &&
      op->et_tiling[ip][2] == tl.ZC)
    return PTMF_OK;
  const long len = tl.ntiles * tl.ZC * 27L * kEPW;
  T *dst = (T *)(sizeof(T) == 8 ? (void *)op->et64 : (void *)op->et32);
  PTMF_CHECK(dst && len <= op->et_len, PTMF_ERR_NOT_SETUP,
             "a11_mult_noalloc: the tile-ordered viscosity of the current tiling is not allocated "
             "(ptmf_op_set_viscosity and ptmf_op_set_tiling build it; a hierarchy's level operator "
             "gets it from its first ptmf_a11_mult)");
  const TileArgs a = tile_args(op);
  long nb = (len + 255) / 256;
  if (nb > 65536) nb = 65536;
  eta_tile_kernel<T><<<dim3((unsigned)nb), dim3(256), 0, op->stream>>>(a, op->d64.eta, dst);
  PTMF_HIP(hipGetLastError());
  op->et_state[ip] = op->state;
  op->et_tiling[ip][0] = tl.TX;
  op->et_tiling[ip][1] = tl.TY;
  op->et_tiling[ip][2] = tl.ZC;
  return PTMF_OK;
}

// The MatMult's kernels with the face buffers at fb (the operator's own, or
// the static scratch of the allocation-free entry point).  noalloc: nothing
// is allocated on the way (the stored-geometry and coloured variants, which
// build tables of their own or need no face buffers, give way to the z-march).
template <typename T>
int launch_a11_mult_fb(ptmf_op_s *op, const T *x, T *y, int faces, T *fb, bool noalloc = false) {
  const OpData<T> &d = *(const OpData<T> *)(sizeof(T) == 8 ? (const void *)&op->d64
                                                           : (const void *)&op->d32);
  const Tiling &tl = op->tiling;
  PTMF_CHECK(tl.TX * tl.TY <= kEPW && (2 * tl.TX + 1) * (2 * tl.TY + 1) <= kWinNodes && tl.ZC > 0,
             PTMF_ERR_ARG, "a11: invalid tiling %d x %d x %d", tl.TX, tl.TY, tl.ZC);
  PTMF_CHECK(fb != nullptr, PTMF_ERR_NOT_SETUP, "a11: face buffers not allocated");
  const int variant = noalloc && op->variant >= 2 ? 0 : op->variant;
  if (variant == 3) {
    // coloured scatter: no tiles, no face buffers
    const Box &b = op->box;
    PTMF_TRY(launch_fill<T>(y, T(0), 3 * b.nnodes, op->stream));
    for (int color = 0; color < 8; color++) {
      const long nb = (class_count(b, color) + kEPW - 1) / kEPW;
      if (nb == 0) continue;
      a11_colour_kernel<T><<<dim3((unsigned)nb), dim3(64), 0, op->stream>>>(
          b.mx, b.my, b.mz, b.nel, b.X, b.X * b.Y, color, d.coords, d.eta, op->bcnode, x, y);
    }
    PTMF_HIP(hipGetLastError());
    if (faces == 1) PTMF_TRY(launch_bc_copy<T>(op->bcidx, op->nbc, x, y, op->stream));
    return PTMF_OK;
  }
  const TileArgs a = tile_args(op);
  const long ntiles = tl.ntiles;
  T *fbx = fb, *fby = fbx + 3 * a.fbx_n * ntiles, *fbz = fby + 3 * a.fby_n * ntiles;
  if (variant == 2) {
    PTMF_TRY(ensure_pa<T>(op));
    const T *pa = sizeof(T) == 8 ? (const T *)op->pa64 : (const T *)op->pa32;
    a11_tile_pa_kernel<T, true><<<dim3((unsigned)ntiles), dim3(64), 0, op->stream>>>(
        a, pa, op->bcnode, x, y, fbx, fby, fbz);
  } else {
    PTMF_TRY(noalloc ? eta_tiled_in_place<T>(op) : ensure_eta_tiled<T>(op));
    const T *et = sizeof(T) == 8 ? (const T *)op->et64 : (const T *)op->et32;
    const int kw = op->wg_tiles;
    const long nwg = (ntiles + kw - 1) / kw;
    if (variant == 1)
      a11_tile_pf_kernel<T, true><<<dim3((unsigned)ntiles), dim3(64), 0, op->stream>>>(
          a, d.coords, et, op->bcnode, x, y, fbx, fby, fbz);
    else if (kw == 2)
      a11_tile_wg_kernel<T, true, 2><<<dim3((unsigned)nwg), dim3(128), 0, op->stream>>>(
          a, nwg, d.coords, et, op->bcnode, x, y, fbx, fby, fbz);
    else if (kw == 4)
      a11_tile_wg_kernel<T, true, 4><<<dim3((unsigned)nwg), dim3(256), 0, op->stream>>>(
          a, nwg, d.coords, et, op->bcnode, x, y, fbx, fby, fbz);
    else
      a11_tile_kernel<T, true><<<dim3((unsigned)ntiles), dim3(64), 0, op->stream>>>(
          a, d.coords, et, op->bcnode, x, y, fbx, fby, fbz);
  }
  PTMF_HIP(hipGetLastError());
  if (faces) {
    a11_face_kernel<T><<<dim3((unsigned)ntiles), dim3(kFaceThreads), 0, op->stream>>>(a, fbx, fby, fbz, y);
    PTMF_HIP(hipGetLastError());
    if (faces == 1) PTMF_TRY(launch_bc_copy<T>(op->bcidx, op->nbc, x, y, op->stream));
  }
  return PTMF_OK;
}

template <typename T>
int launch_a11_mult(ptmf_op_s *op, const T *x, T *y, int faces) {
  return launch_a11_mult_fb<T>(op, x, y, faces, (T *)op->faces);
}

template <typename T>
int launch_stokes_mult_t(ptmf_op_s *op, const T *x, T *y) {
  const OpData<T> &d = *(const OpData<T> *)(sizeof(T) == 8 ? (const void *)&op->d64
                                                           : (const void *)&op->d32);
  const Tiling &tl = op->tiling;
  PTMF_CHECK(tl.TX * tl.TY <= kEPW && (2 * tl.TX + 1) * (2 * tl.TY + 1) <= kWinNodes && tl.ZC > 0,
             PTMF_ERR_ARG, "stokes: invalid tiling %d x %d x %d", tl.TX, tl.TY, tl.ZC);
  PTMF_CHECK(op->faces != nullptr, PTMF_ERR_NOT_SETUP, "stokes: face buffers not allocated");
  PTMF_TRY(ensure_eta_tiled<T>(op));
  const T *et = sizeof(T) == 8 ? (const T *)op->et64 : (const T *)op->et32;
  const TileArgs a = tile_args(op);
  const long ntiles = tl.ntiles, nu = 3 * op->box.nnodes;
  T *fbx = (T *)op->faces, *fby = fbx + 3 * a.fbx_n * ntiles, *fbz = fby + 3 * a.fby_n * ntiles;
  stokes_tile_kernel<T><<<dim3((unsigned)ntiles), dim3(64), 0, op->stream>>>(
      a, d.coords, et, op->bcnode, x, x + nu, y, y + nu, fbx, fby, fbz);
  PTMF_HIP(hipGetLastError());
  a11_face_kernel<T><<<dim3((unsigned)ntiles), dim3(kFaceThreads), 0, op->stream>>>(a, fbx, fby, fbz, y);
  PTMF_HIP(hipGetLastError());
  PTMF_TRY(launch_bc_copy<T>(op->bcidx, op->nbc, x, y, op->stream));
  return PTMF_OK;
}

int launch_stokes_mult(ptmf_op_s *op, const double *x, double *y) { return launch_stokes_mult_t<double>(op, x, y); }

// FP32 saddle-point MatMult (FP32 coordinates, viscosity and accumulation):
// the preconditioner's A12 product of an FP32 velocity V-cycle.
int launch_stokes_mult_f32(ptmf_op_s *op, const float *x, float *y) { return launch_stokes_mult_t<float>(op, x, y); }

int prepare_eta_tiled(ptmf_op_s *op) { return ensure_eta_tiled<double>(op); }

template int launch_a11<double>(const ptmf_op_s *, const double *, double *, bool);
template int launch_a11<float>(const ptmf_op_s *, const float *, float *, bool);
template int launch_a11_mult<double>(ptmf_op_s *, const double *, double *, int);
template int launch_a11_mult<float>(ptmf_op_s *, const float *, float *, int);

}  // namespace ptmf

// ---------------------------------------------------------------------------
// Batched MatMult object (C-ABI)
// ---------------------------------------------------------------------------
struct ptmf_batch_s {
  std::vector<ptmf_op> ops;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  // device data, rebuilt when the operators' tilings or buffers change
  ptmf::BoxDesc *dbox[2] = {nullptr, nullptr};
  std::vector<unsigned char> hbox[2];
  int2 *dwork = nullptr;
  long nwork = 0;
  std::vector<int> key;  // TX, TY, ZC, ntiles per operator of the current work list
  void **dxy[2] = {nullptr, nullptr};  // device arrays of x and y pointers, per precision
  std::vector<const void *> hxy[2];
  long maxnbc = 0;
};

namespace {

// work list and per-box descriptors of precision T (no vectors)
template <typename T>
int batch_prepare_boxes(ptmf_batch_s *b) {
  const int ip = sizeof(T) == 8 ? 0 : 1;
  const int n = (int)b->ops.size();
  // work list of (box, tile)
  std::vector<int> key;
  for (ptmf_op op : b->ops) {
    const ptmf::Tiling &t = op->tiling;
    key.insert(key.end(), {t.TX, t.TY, t.ZC, (int)t.ntiles});
  }
  if (key != b->key) {
    std::vector<int2> work;
    for (int i = 0; i < n; i++)
      for (long t = 0; t < b->ops[i]->tiling.ntiles; t++) work.push_back(make_int2(i, (int)t));
    PTMF_CHECK(!work.empty() && work.size() < (1UL << 31), PTMF_ERR_ARG, "batch: %zu tiles", work.size());
    (void)hipFree(b->dwork);
    b->dwork = nullptr;
    PTMF_HIP(hipMalloc((void **)&b->dwork, sizeof(int2) * work.size()));
    PTMF_HIP(hipMemcpy(b->dwork, work.data(), sizeof(int2) * work.size(), hipMemcpyHostToDevice));
    b->nwork = (long)work.size();
    b->key = key;
  }
  // per-box descriptors of this precision
  std::vector<ptmf::BoxDesc> box(n);
  b->maxnbc = 0;
  for (int i = 0; i < n; i++) {
    ptmf_op op = b->ops[i];
    PTMF_CHECK(op->has_coords && op->has_eta && op->has_bc, PTMF_ERR_NOT_SETUP,
               "batch: operator %d: MF operator not setup", i);
    PTMF_TRY(ptmf::ensure_eta_tiled<T>(op));
    ptmf::BoxDesc &d = box[i];
    d.a = ptmf::tile_args(op);
    const ptmf::OpData<T> &od = *(const ptmf::OpData<T> *)(ip == 0 ? (const void *)&op->d64 : (const void *)&op->d32);
    d.xc = od.coords;
    d.eta = ip == 0 ? (const void *)op->et64 : (const void *)op->et32;
    d.bcn = op->bcnode;
    d.bcidx = op->bcidx;
    d.nbc = op->nbc;
    T *fb = (T *)op->faces;
    d.fbx = fb;
    d.fby = fb + 3 * d.a.fbx_n * d.a.ntiles;
    d.fbz = (T *)d.fby + 3 * d.a.fby_n * d.a.ntiles;
    if (op->nbc > b->maxnbc) b->maxnbc = op->nbc;
  }
  // the tile-ordered viscosity is built on each operator's own stream
  for (ptmf_op op : b->ops) PTMF_HIP(hipStreamSynchronize(op->stream));
  const size_t bytes = sizeof(ptmf::BoxDesc) * n;
  if (!b->dbox[ip]) PTMF_HIP(hipMalloc((void **)&b->dbox[ip], bytes));
  const unsigned char *raw = (const unsigned char *)box.data();
  if (b->hbox[ip].size() != bytes || memcmp(b->hbox[ip].data(), raw, bytes) != 0) {
    b->hbox[ip].assign(raw, raw + bytes);
    PTMF_HIP(hipMemcpyAsync(b->dbox[ip], b->hbox[ip].data(), bytes, hipMemcpyHostToDevice, b->stream));
  }
  return PTMF_OK;
}

template <typename T>
int batch_prepare(ptmf_batch_s *b, const void *const *x, void *const *y) {
  const int ip = sizeof(T) == 8 ? 0 : 1;
  const int n = (int)b->ops.size();
  PTMF_TRY(batch_prepare_boxes<T>(b));
  // x and y pointer arrays
  std::vector<const void *> xy(2 * n);
  for (int i = 0; i < n; i++) {
    PTMF_CHECK(x[i] && y[i] && x[i] != y[i], PTMF_ERR_ARG, "batch: vectors of box %d", i);
    xy[i] = x[i];
    xy[n + i] = y[i];
  }
  if (!b->dxy[ip]) PTMF_HIP(hipMalloc((void **)&b->dxy[ip], sizeof(void *) * 2 * n));
  if (b->hxy[ip] != xy) {
    b->hxy[ip] = xy;
    PTMF_HIP(hipMemcpyAsync(b->dxy[ip], b->hxy[ip].data(), sizeof(void *) * 2 * n, hipMemcpyHostToDevice,
                            b->stream));
  }
  return PTMF_OK;
}

// the three batched stages on device tables of x and y pointers
template <typename T>
int batch_launch_tables(ptmf_batch_s *b, const T *const *xs, T *const *ys, hipStream_t s,
                        const int *active = nullptr) {
  const int ip = sizeof(T) == 8 ? 0 : 1;
  const int n = (int)b->ops.size();
  ptmf::a11_tile_batch_kernel<T><<<dim3((unsigned)b->nwork), dim3(64), 0, s>>>(b->dbox[ip], b->dwork, b->nwork, xs,
                                                                              ys, active);
  PTMF_HIP(hipGetLastError());
  ptmf::a11_face_batch_kernel<T><<<dim3((unsigned)b->nwork), dim3(64), 0, s>>>(b->dbox[ip], b->dwork, b->nwork, ys,
                                                                              active);
  PTMF_HIP(hipGetLastError());
  if (b->maxnbc > 0) {
    long gx = (b->maxnbc + 255) / 256;
    if (gx > 1024) gx = 1024;
    ptmf::bc_copy_batch_kernel<T><<<dim3((unsigned)gx, (unsigned)n), dim3(256), 0, s>>>(b->dbox[ip], xs, ys,
                                                                                        active);
    PTMF_HIP(hipGetLastError());
  }
  return PTMF_OK;
}

template <typename T>
int batch_launch(ptmf_batch_s *b) {
  const int ip = sizeof(T) == 8 ? 0 : 1;
  const int n = (int)b->ops.size();
  return batch_launch_tables<T>(b, (const T *const *)b->dxy[ip], (T *const *)(b->dxy[ip] + n), b->stream);
}

}  // namespace

namespace ptmf {
int batch_bind(ptmf_batch_s *b, int precision) {
  for (ptmf_op op : b->ops) PTMF_HIP(hipStreamSynchronize(op->stream));
  if (precision == 64) PTMF_TRY(batch_prepare_boxes<double>(b));
  else PTMF_TRY(batch_prepare_boxes<float>(b));
  PTMF_HIP(hipStreamSynchronize(b->stream));
  return PTMF_OK;
}

int batch_apply(ptmf_batch_s *b, int precision, const void *const *xs, void *const *ys, hipStream_t s,
                const int *active) {
  if (precision == 64)
    return batch_launch_tables<double>(b, (const double *const *)xs, (double *const *)ys, s, active);
  return batch_launch_tables<float>(b, (const float *const *)xs, (float *const *)ys, s, active);
}
}  // namespace ptmf

extern "C" {

int ptmf_batch_create(const ptmf_op *ops, int nops, ptmf_batch *out) {
  PTMF_CHECK(ops && out && nops > 0 && nops <= 65535, PTMF_ERR_ARG, "batch_create: bad arguments");
  for (int i = 0; i < nops; i++) PTMF_CHECK(ops[i], PTMF_ERR_ARG, "batch_create: null operator %d", i);
  ptmf_batch_s *b = new ptmf_batch_s;
  b->ops.assign(ops, ops + nops);
  hipError_t e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreate(&b->ev0);
  if (e == hipSuccess) e = hipEventCreate(&b->ev1);
  if (e != hipSuccess) {
    ptmf_batch tmp = b;
    ptmf_batch_destroy(&tmp);
    return ptmf_set_error(PTMF_ERR_DEVICE, "batch_create: %s", hipGetErrorString(e));
  }
  *out = b;
  return PTMF_OK;
}

int ptmf_batch_destroy(ptmf_batch *pb) {
  if (!pb || !*pb) return PTMF_OK;
  ptmf_batch_s *b = *pb;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  for (int i = 0; i < 2; i++) {
    (void)hipFree(b->dbox[i]);
    (void)hipFree(b->dxy[i]);
  }
  (void)hipFree(b->dwork);
  if (b->ev0) (void)hipEventDestroy(b->ev0);
  if (b->ev1) (void)hipEventDestroy(b->ev1);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  *pb = nullptr;
  return PTMF_OK;
}

int ptmf_batch_mult(ptmf_batch b, int precision, const void *const *x, void *const *y) {
  PTMF_CHECK(b && x && y, PTMF_ERR_ARG, "batch_mult: null argument");
  PTMF_CHECK(precision == 64 || precision == 32, PTMF_ERR_UNSUPPORTED, "batch_mult: precision %d", precision);
  for (ptmf_op op : b->ops) PTMF_HIP(hipStreamSynchronize(op->stream));
  if (precision == 64) {
    PTMF_TRY(batch_prepare<double>(b, x, y));
    PTMF_TRY(batch_launch<double>(b));
  } else {
    PTMF_TRY(batch_prepare<float>(b, x, y));
    PTMF_TRY(batch_launch<float>(b));
  }
  PTMF_HIP(hipStreamSynchronize(b->stream));
  return PTMF_OK;
}

int ptmf_batch_time(ptmf_batch b, int precision, const void *const *x, void *const *y, int reps, double *ms) {
  PTMF_CHECK(b && x && y && reps > 0 && ms, PTMF_ERR_ARG, "batch_time: bad arguments");
  PTMF_TRY(ptmf_batch_mult(b, precision, x, y));  // prepares and warms up
  PTMF_HIP(hipEventRecord(b->ev0, b->stream));
  for (int r = 0; r < reps; r++) {
    if (precision == 64) PTMF_TRY(batch_launch<double>(b));
    else PTMF_TRY(batch_launch<float>(b));
  }
  PTMF_HIP(hipEventRecord(b->ev1, b->stream));
  PTMF_HIP(hipEventSynchronize(b->ev1));
  float t = 0.0f;
  PTMF_HIP(hipEventElapsedTime(&t, b->ev0, b->ev1));
  *ms = (double)t / reps;
  return PTMF_OK;
}

}  // extern "C"


// ---------------------------------------------------------------------------
// Allocation-free MatMult (ptmf_a11_mult_noalloc): the FP64 MatMult of
// ptmf_a11_mult with the z-march's face buffers in `scratch`, a file-scope
// device array of kScratchDoubles doubles (64 MiB) that every GPU the library
// runs on holds until the process ends, instead of the operator's own
// allocation.  In host code `scratch` names a host-side stand-in, not device
// memory: the device address comes from hipGetSymbolAddress, once per
// device, with that device current, and only that address reaches kernels.
// Every call on one GPU shares its array (the tile kernel writes every face
// entry the face kernel reads, so nothing needs clearing): calls that overlap
// on one GPU race.
// ---------------------------------------------------------------------------
#include <mutex>

namespace {
constexpr long kScratchDoubles = 8L << 20;
constexpr int kScratchDevices = 64;
}  // namespace

static __device__ __align__(256) double scratch[kScratchDoubles];

namespace {
std::mutex scratch_mu;
double *scratch_at[kScratchDevices];

// device address of `scratch` on device dev (the current device)
int scratch_address(int dev, double **out) {
  PTMF_CHECK(dev >= 0 && dev < kScratchDevices, PTMF_ERR_DEVICE, "a11_mult_noalloc: device %d", dev);
  std::lock_guard<std::mutex> lock(scratch_mu);
  if (!scratch_at[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    PTMF_CHECK(e == hipSuccess && p, PTMF_ERR_DEVICE, "a11_mult_noalloc: hipGetSymbolAddress(scratch): %s",
               hipGetErrorString(e));
    scratch_at[dev] = (double *)p;
  }
  *out = scratch_at[dev];
  return PTMF_OK;
}
}  // namespace

extern "C" int ptmf_a11_mult_noalloc(ptmf_op op, const double *x, double *y) {
  PTMF_CHECK(op, PTMF_ERR_ARG, "a11_mult_noalloc: null operator");
  PTMF_CHECK(op->has_coords && op->has_eta && op->has_bc, PTMF_ERR_NOT_SETUP, "MF operator not setup");
  PTMF_CHECK(x && y && (const void *)x != (const void *)y, PTMF_ERR_ARG,
             "a11_mult_noalloc: x and y must be distinct device arrays");
  int dev = -1;
  PTMF_HIP(hipGetDevice(&dev));
  PTMF_CHECK(dev == op->device, PTMF_ERR_ARG,
             "a11_mult_noalloc: the operator lives on device %d but device %d is current", op->device, dev);
  const ptmf::Tiling &tl = op->tiling;
  const long need = ptmf::face_buffer_doubles(tl.TX, tl.TY, tl.ZC, tl.ntiles);
  PTMF_CHECK(need <= kScratchDoubles, PTMF_ERR_UNSUPPORTED,
             "a11_mult_noalloc: the face buffers of this box and tiling take %ld doubles, the static scratch "
             "holds %ld (64 MiB); use ptmf_a11_mult",
             need, kScratchDoubles);
  double *fb = nullptr;
  PTMF_TRY(scratch_address(dev, &fb));
  return ptmf::launch_a11_mult_fb<double>(op, x, y, 1, fb, true);
}
