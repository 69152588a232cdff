// a11.hip -- matrix-free Q2 viscous block (A11) apply for gfx950.
//
// Computes y += A11_local u for all elements of a structured Q2 box, the
// operator of MFStokesWrapper_A11_Tensor (stokes_operators_tensor.c:181-288):
// per element and quadrature point, grad u and the Jacobian from the 27
// nodal values, then w*detJ*(2 eta sym(grad u)) G^T, then the transposed
// gradient back to the nodes.
//
// Work decomposition (MI355X-first, not the reference's one-warp-per-element):
//   * 3 lanes per element, lane c owns quadrature plane zeta_c (9 qps) and,
//     on input/output, node plane k = c of the element; 21 elements per
//     64-lane wave, one wave per work-group (no cross-wave barriers).
//   * Collocated sum factorisation: nodal -> quadrature-grid interpolation
//     (B x B x B), then the three reference derivatives on the quadrature
//     grid with the 3x3 collocation matrix C = D B^-1.  The xi/eta stages
//     are in registers; only the zeta stage crosses lanes, through LDS
//     (9 values per field per lane instead of 27).
//   * The 1/detJ of the inverse Jacobian is folded into one scalar per qp
//     (adjugate form), so per qp there is one division.
//   * Scatter: hardware FP64/FP32 atomic add into y (global_atomic_add_f64);
//     y is cleared beforehand by the caller.
//   * Viscosity is stored quadrature-point-major (eta[q*nel + e]) so the
//     lanes of a wave read consecutive elements.
#include "ptmf_internal.h"

namespace ptmf {
namespace {

constexpr int kEPW = 21;  // elements per wave (3 lanes each; lane 63 idle)
constexpr int kSlots = 54;

template <typename T>
__device__ __forceinline__ T sel3(int c, double a0, double a1, double a2) {
  return (T)(c == 0 ? a0 : (c == 1 ? a1 : a2));
}

// Load node plane (base + i + X*j, i,j = 0..2) of a 3-component field and
// interpolate it to the 3x3 in-plane quadrature grid: V[d][3b+a] =
// sum_{i,j} B[a][i] B[b][j] f_d(i,j).
template <typename T, bool kMask>
__device__ __forceinline__ void load_interp_plane(const T *__restrict__ f, long base, long X,
                                                  const uint8_t *__restrict__ bcn, T V[3][9]) {
  T g[3][3][3];  // [d][j][i]
#pragma unroll
  for (int j = 0; j < 3; j++)
#pragma unroll
    for (int i = 0; i < 3; i++) {
      const long node = base + i + X * j;
      unsigned m = 0;
      if (kMask) m = bcn[node];
#pragma unroll
      for (int d = 0; d < 3; d++) {
        T v = f[3 * node + d];
        g[d][j][i] = ((m >> d) & 1u) ? T(0) : v;
      }
    }
#pragma unroll
  for (int d = 0; d < 3; d++) {
    T t[3][3];  // [j][a]
#pragma unroll
    for (int j = 0; j < 3; j++)
#pragma unroll
      for (int a = 0; a < 3; a++)
        t[j][a] = (T)kB[a][0] * g[d][j][0] + (T)kB[a][1] * g[d][j][1] + (T)kB[a][2] * g[d][j][2];
#pragma unroll
    for (int b = 0; b < 3; b++)
#pragma unroll
      for (int a = 0; a < 3; a++)
        V[d][3 * b + a] = (T)kB[b][0] * t[0][a] + (T)kB[b][1] * t[1][a] + (T)kB[b][2] * t[2][a];
  }
}

template <typename T, bool kMask>
__global__ void __launch_bounds__(64)
    a11_apply_kernel(int mx, int my, long nel, long X, long XY, const T *__restrict__ xc,
                     const T *__restrict__ eta, const uint8_t *__restrict__ bcn,
                     const T *__restrict__ u, T *__restrict__ y) {
  __shared__ T sh[3][kSlots][kEPW];

  const int lane = threadIdx.x;
  const int c = lane / kEPW;  // quadrature / node plane; 3 marks the idle lane
  const int le = lane - c * kEPW;
  const int cp = c < 3 ? c : 2;
  long e = (long)blockIdx.x * kEPW + le;
  const bool live = (c < 3) && (e < nel);
  if (e >= nel) e = nel - 1;
  const bool writer = c < 3;

  const long ei = e % mx, er = e / mx, ej = er % my, ek = er / my;
  const long base = 2 * ei + X * (2 * ej) + XY * (2 * ek + cp);

  // zeta tables of this lane's plane: forward B[c][k], D[c][k]; backward
  // (transposed) B[c'][k=c], D[c'][k=c].
  T bz[3], dz[3], bt[3], dt[3];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    bz[k] = sel3<T>(cp, kB[0][k], kB[1][k], kB[2][k]);
    dz[k] = sel3<T>(cp, kD[0][k], kD[1][k], kD[2][k]);
    bt[k] = sel3<T>(cp, kB[k][0], kB[k][1], kB[k][2]);
    dt[k] = sel3<T>(cp, kD[k][0], kD[k][1], kD[k][2]);
  }
  const T wz = sel3<T>(cp, kW[0], kW[1], kW[2]);

  // ---- forward: coordinates -> Xq (values on the qp plane), Xz (d/dzeta)
  T Xq[3][9], Xz[3][9], Uq[3][9], Uz[3][9];
  {
    T V[3][9];
    load_interp_plane<T, false>(xc, base, X, nullptr, V);
    if (writer) {
#pragma unroll
      for (int d = 0; d < 3; d++)
#pragma unroll
        for (int p = 0; p < 9; p++) sh[cp][9 * d + p][le] = V[d][p];
    }
    __syncthreads();
#pragma unroll
    for (int d = 0; d < 3; d++)
#pragma unroll
      for (int p = 0; p < 9; p++) {
        const T v0 = sh[0][9 * d + p][le], v1 = sh[1][9 * d + p][le], v2 = sh[2][9 * d + p][le];
        Xq[d][p] = bz[0] * v0 + bz[1] * v1 + bz[2] * v2;
        Xz[d][p] = dz[0] * v0 + dz[1] * v1 + dz[2] * v2;
      }
  }
  // ---- forward: velocity (Dirichlet columns masked) -> Uq, Uz
  {
    T V[3][9];
    load_interp_plane<T, kMask>(u, base, X, bcn, V);
    if (writer) {
#pragma unroll
      for (int d = 0; d < 3; d++)
#pragma unroll
        for (int p = 0; p < 9; p++) sh[cp][27 + 9 * d + p][le] = V[d][p];
    }
    __syncthreads();
#pragma unroll
    for (int d = 0; d < 3; d++)
#pragma unroll
      for (int p = 0; p < 9; p++) {
        const T v0 = sh[0][27 + 9 * d + p][le], v1 = sh[1][27 + 9 * d + p][le],
                v2 = sh[2][27 + 9 * d + p][le];
        Uq[d][p] = bz[0] * v0 + bz[1] * v1 + bz[2] * v2;
        Uz[d][p] = dz[0] * v0 + dz[1] * v1 + dz[2] * v2;
      }
  }

  // ---- pointwise action on the 9 qps of this plane
  T W[3][9];  // in-plane transposed xi/eta derivatives, accumulated
#pragma unroll
  for (int l = 0; l < 3; l++)
#pragma unroll
    for (int p = 0; p < 9; p++) W[l][p] = T(0);

  const T *etap = eta + (long)(9 * cp) * nel + e;
#pragma unroll
  for (int b = 0; b < 3; b++)
#pragma unroll
    for (int a = 0; a < 3; a++) {
      const int p = 3 * b + a;
      // M[j][d] = d x_d / d xi_j, Gu[j][l] = d u_l / d xi_j
      T M[3][3], Gu[3][3];
#pragma unroll
      for (int d = 0; d < 3; d++) {
        M[0][d] = (T)kC[a][0] * Xq[d][3 * b] + (T)kC[a][1] * Xq[d][3 * b + 1] +
                  (T)kC[a][2] * Xq[d][3 * b + 2];
        M[1][d] = (T)kC[b][0] * Xq[d][a] + (T)kC[b][1] * Xq[d][3 + a] + (T)kC[b][2] * Xq[d][6 + a];
        M[2][d] = Xz[d][p];
        Gu[0][d] = (T)kC[a][0] * Uq[d][3 * b] + (T)kC[a][1] * Uq[d][3 * b + 1] +
                   (T)kC[a][2] * Uq[d][3 * b + 2];
        Gu[1][d] = (T)kC[b][0] * Uq[d][a] + (T)kC[b][1] * Uq[d][3 + a] + (T)kC[b][2] * Uq[d][6 + a];
        Gu[2][d] = Uz[d][p];
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
      const T s = (T(2) * wq * etap[(long)p * nel]) / det;
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
      // dv[j][l] = w detJ sum_m sigma[m][l] d xi_j / d x_m
#pragma unroll
      for (int l = 0; l < 3; l++) {
        const T dv0 = sg[0][l] * A[0][0] + sg[1][l] * A[1][0] + sg[2][l] * A[2][0];
        const T dv1 = sg[0][l] * A[0][1] + sg[1][l] * A[1][1] + sg[2][l] * A[2][1];
        const T dv2 = sg[0][l] * A[0][2] + sg[1][l] * A[1][2] + sg[2][l] * A[2][2];
#pragma unroll
        for (int r = 0; r < 3; r++) {
          W[l][3 * b + r] += (T)kC[a][r] * dv0;
          W[l][3 * r + a] += (T)kC[b][r] * dv1;
        }
        if (writer) sh[cp][9 * l + p][le] = dv2;  // zeta component, exchanged below
      }
    }
  __syncthreads();
  if (writer) {
#pragma unroll
    for (int l = 0; l < 3; l++)
#pragma unroll
      for (int p = 0; p < 9; p++) sh[cp][27 + 9 * l + p][le] = W[l][p];
  }
  __syncthreads();

  // ---- backward zeta stage: node plane k = c gathers all three qp planes
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

  // ---- backward in-plane: transposed interpolation, then scatter-add
#pragma unroll
  for (int l = 0; l < 3; l++) {
    T t[3][3];  // [b][i]
#pragma unroll
    for (int b = 0; b < 3; b++)
#pragma unroll
      for (int i = 0; i < 3; i++)
        t[b][i] = (T)kB[0][i] * Y[l][3 * b] + (T)kB[1][i] * Y[l][3 * b + 1] +
                  (T)kB[2][i] * Y[l][3 * b + 2];
#pragma unroll
    for (int j = 0; j < 3; j++)
#pragma unroll
      for (int i = 0; i < 3; i++) {
        const T v = (T)kB[0][j] * t[0][i] + (T)kB[1][j] * t[1][i] + (T)kB[2][j] * t[2][i];
        if (live) atomicAdd(&y[3 * (base + i + X * j) + l], v);
      }
  }
}

}  // namespace

template <typename T>
int launch_a11(const ptmf_op_s *op, const T *u, T *y, bool mask_columns) {
  const Box &b = op->box;
  const OpData<T> &d = *(const OpData<T> *)(sizeof(T) == 8 ? (const void *)&op->d64
                                                           : (const void *)&op->d32);
  const long nblk = (b.nel + kEPW - 1) / kEPW;
  PTMF_CHECK(nblk < (1L << 31), PTMF_ERR_ARG, "a11: mesh too large (%ld blocks)", nblk);
  if (mask_columns)
    a11_apply_kernel<T, true><<<dim3((unsigned)nblk), dim3(64), 0, op->stream>>>(
        b.mx, b.my, b.nel, b.X, b.X * b.Y, d.coords, d.eta, op->bcnode, u, y);
  else
    a11_apply_kernel<T, false><<<dim3((unsigned)nblk), dim3(64), 0, op->stream>>>(
        b.mx, b.my, b.nel, b.X, b.X * b.Y, d.coords, d.eta, nullptr, u, y);
  PTMF_HIP(hipGetLastError());
  return PTMF_OK;
}

template int launch_a11<double>(const ptmf_op_s *, const double *, double *, bool);
template int launch_a11<float>(const ptmf_op_s *, const float *, float *, bool);

}  // namespace ptmf
