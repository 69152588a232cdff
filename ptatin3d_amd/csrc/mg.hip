// mg.hip -- geometric multigrid on the matrix-free A11 operator: transfers,
// Chebyshev-Jacobi smoothing, coarse solve, V-cycle and FP64 FGMRES.
//
// Mirrors the reference's viscous-block preconditioner
// (pTatin3dStokesKSPConfigureFSGMG, ptatin_driver_linear_ts.c:819-882):
// a PCMG V-cycle whose levels are matrix-free A11 operators, KSPCHEBYSHEV +
// PCJACOBI smoothers (recurrence of ksp_chebyrn.c:322-411), DMDA Q1 lattice
// interpolation P with restriction P^T (ptatin_driver_linear_ts.c:987-996),
// and an exact coarse solve.  Level 0 is the coarsest, as in PCMG.
//
// MI355X choices:
//   * every vector stays in HBM for the whole solve; the host only sees the
//     Krylov scalars;
//   * the V-cycle runs in FP32 (operator, smoother, transfers) or FP64; the
//     outer FGMRES, its true residual and the coarse solve are FP64;
//   * restriction is written in gather form (one thread per coarse node), so
//     neither transfer needs atomics;
//   * each Chebyshev step is one MatMult plus one fused update kernel;
//   * the coarsest level is factorised once (dense Cholesky, rocSOLVER).
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include "ptmf_internal.h"

namespace ptmf {
namespace {

constexpr int kBS = 256;
constexpr int kRedBlocks = 1024;
constexpr int kMaxDots = 512;
constexpr int kRedSlots = 2 * kMaxDots + 4;  // reduction results per synchronisation
constexpr long kDenseMaxDofs = 32768;  // largest level 0 of the dense coarse inverse

inline unsigned nblocks(long n) {
  long g = (n + kBS - 1) / kBS;
  if (g > 65536) g = 65536;
  if (g < 1) g = 1;
  return (unsigned)g;
}

// ---------------------------------------------------------------------------
// vector kernels
// ---------------------------------------------------------------------------
template <typename T>
__global__ void k_axpby(long n, T a, const T *__restrict__ x, T b, T *__restrict__ y) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    y[i] = a * x[i] + b * y[i];
}

// r = b - Ax; r may alias Ax (elementwise, same index)
template <typename T>
__global__ void k_residual(long n, const T *__restrict__ b, const T *Ax, T *r) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    r[i] = b[i] - Ax[i];
}

// r = b - Ax with an FP32 product
__global__ void k_residual_mixed(long n, const double *__restrict__ b, const float *__restrict__ Ax,
                                 double *__restrict__ r) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    r[i] = b[i] - (double)Ax[i];
}

template <typename Ts, typename Td>
__global__ void k_cast(long n, const Ts *__restrict__ x, Td *__restrict__ y) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    y[i] = (Td)x[i];
}

template <typename T>
__global__ void k_copy(long n, const T *__restrict__ x, T *__restrict__ y) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    y[i] = x[i];
}

// Dense column-major product y = A x (A: n x m, leading dimension lda) in
// two launches: each work-group forms the partial sums of kGvRows rows over
// one chunk of kGvCols columns (x chunk staged in LDS, consecutive threads
// read consecutive rows of a column), then the chunks are summed in a fixed
// order, so the result is deterministic and no atomics are needed.
constexpr int kGvRows = 256;
constexpr int kGvCols = 512;

__global__ void __launch_bounds__(kGvRows)
    k_gemv_part(long n, long m, const double *__restrict__ A, long lda, const double *__restrict__ x,
                double *__restrict__ part) {
  __shared__ double xs[kGvCols];
  const long j0 = (long)blockIdx.y * kGvCols;
  const int nj = (int)(m - j0 < kGvCols ? m - j0 : kGvCols);
  for (int t = threadIdx.x; t < nj; t += kGvRows) xs[t] = x[j0 + t];
  __syncthreads();
  const long i = (long)blockIdx.x * kGvRows + threadIdx.x;
  if (i >= n) return;
  const double *a = A + i + j0 * lda;
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
  int j = 0;
  for (; j + 4 <= nj; j += 4) {
    s0 += a[(long)j * lda] * xs[j];
    s1 += a[(long)(j + 1) * lda] * xs[j + 1];
    s2 += a[(long)(j + 2) * lda] * xs[j + 2];
    s3 += a[(long)(j + 3) * lda] * xs[j + 3];
  }
  for (; j < nj; j++) s0 += a[(long)j * lda] * xs[j];
  part[(long)blockIdx.y * n + i] = (s0 + s1) + (s2 + s3);
}

__global__ void k_gemv_sum(long n, int nchunks, const double *__restrict__ part, double *__restrict__ y) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    double s = 0.0;
    for (int c = 0; c < nchunks; c++) s += part[(long)c * n + i];
    y[i] = s;
  }
}

// mirror the lower triangle of a column-major n x n matrix into its upper one
__global__ void k_symmetrize(long n, double *__restrict__ A) {
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n * n; t += (long)gridDim.x * blockDim.x) {
    const long i = t % n, j = t / n;
    if (i < j) A[i + j * n] = A[j + i * n];
  }
}

// y += (Td) x
template <typename Ts, typename Td>
__global__ void k_add_cast(long n, const Ts *__restrict__ x, Td *__restrict__ y) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    y[i] += (Td)x[i];
}

template <typename T>
__global__ void k_recip(long n, const double *__restrict__ d, T *__restrict__ dinv) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    dinv[i] = (T)(1.0 / d[i]);
}

// first Chebyshev step: p = x + scale * dinv * (b - Ax)   (Ax == nullptr: x = 0)
// b may be of a wider type Tb than T (the FP64 right-hand side of an FP32
// V-cycle): it is rounded to T first and, with bout, kept in T for the
// later steps (the cast and the first step in one pass).
template <typename T, typename Tb = T>
__global__ void k_cheb_first(long n, const Tb *__restrict__ b, const T *__restrict__ Ax,
                             const T *__restrict__ x, const T *__restrict__ dinv, T scale,
                             T *__restrict__ p, T *__restrict__ bout = nullptr) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const T bi = (T)b[i];
    if (bout) bout[i] = bi;
    const T r = Ax ? bi - Ax[i] : bi;
    const T x0 = x ? x[i] : T(0);
    p[i] = x0 + scale * (r * dinv[i]);
  }
}

// p_kp1 = (omega scale) dinv (b - A p_k) + (1 - omega) p_km1 + omega p_k;
// with bcn (per-node Dirichlet bits) the Dirichlet rows of A p_k are taken as
// p_k (the MatMult's y = x), so Apk may hold the element sums there
// pkp1 may be pkm1 (the step overwrites the iterate before last in place);
// pkm1 == nullptr stands for a zero iterate.  To: the type the new iterate
// is stored in (the FP64 output of an FP32 V-cycle's last step: the step
// and the cast in one pass).
template <typename T, typename To = T>
__global__ void k_cheb_step(long n, const T *__restrict__ b, const T *__restrict__ Apk,
                            const T *__restrict__ dinv, const T *pkm1,
                            const T *__restrict__ pk, T os, T om1, T om, To *pkp1,
                            const uint8_t *__restrict__ bcn = nullptr) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    T ap = Apk[i];
    if (bcn) {
      const long nd = i / 3;
      if ((bcn[nd] >> (int)(i - 3 * nd)) & 1u) ap = pk[i];
    }
    const T z = (b[i] - ap) * dinv[i];
    const T prev = pkm1 ? pkm1[i] : T(0);
    pkp1[i] = (To)(os * z + om1 * prev + om * pk[i]);
  }
}

// partial sums of the dot products <V_j, w>, j in [j0, j0 + nk), nk <= 8,
// one value per block: part[j * gridDim.x + block]
__global__ void k_mdot_partial(long n, int j0, int nk, const double *const *__restrict__ V,
                               const double *__restrict__ w, double *__restrict__ part) {
  __shared__ double red[8][kBS / 64];
  double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const double wi = w[i];
#pragma unroll
    for (int j = 0; j < 8; j++)
      if (j < nk) acc[j] += V[j0 + j][i] * wi;
  }
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int j = 0; j < 8; j++) {
    double v = acc[j];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    if (lane == 0) red[j][wv] = v;
  }
  __syncthreads();
  if (threadIdx.x < nk) {
    double t = 0.0;
    for (int q = 0; q < kBS / 64; q++) t += red[threadIdx.x][q];
    part[(long)(j0 + threadIdx.x) * gridDim.x + blockIdx.x] = t;
  }
}

// partial sums of r.r, z.r and z.q: part[j * gridDim.x + block]
__global__ void k_dot3_partial(long n, const double *__restrict__ r, const double *__restrict__ z,
                               const double *__restrict__ q, double *__restrict__ part) {
  __shared__ double red[3][kBS / 64];
  double a0 = 0.0, a1 = 0.0, a2 = 0.0;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const double ri = r[i], zi = z[i];
    a0 += ri * ri;
    a1 += zi * ri;
    a2 += zi * q[i];
  }
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int off = 32; off > 0; off >>= 1) {
    a0 += __shfl_down(a0, off, 64);
    a1 += __shfl_down(a1, off, 64);
    a2 += __shfl_down(a2, off, 64);
  }
  if (lane == 0) {
    red[0][wv] = a0;
    red[1][wv] = a1;
    red[2][wv] = a2;
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    double t = 0.0;
    for (int w = 0; w < kBS / 64; w++) t += red[threadIdx.x][w];
    part[(long)threadIdx.x * gridDim.x + blockIdx.x] = t;
  }
}

// partial sums of one dot product <a, b>
// the flexible-CG step x += alpha p, r -= alpha q in one pass
__global__ void k_fcg_step(long n, double alpha, const double *__restrict__ p, const double *__restrict__ q,
                           double *__restrict__ x, double *__restrict__ r) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    x[i] += alpha * p[i];
    r[i] -= alpha * q[i];
  }
}

__global__ void k_dot_partial(long n, const double *__restrict__ a, const double *__restrict__ b,
                              double *__restrict__ part) {
  __shared__ double red[kBS / 64];
  double acc = 0.0;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    acc += a[i] * b[i];
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int q = 0; q < kBS / 64; q++) t += red[q];
    part[blockIdx.x] = t;
  }
}

// y = a x
template <typename T>
__global__ void k_scale(long n, T a, const T *__restrict__ x, T *__restrict__ y) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    y[i] = a * x[i];
}

__global__ void k_mdot_final(int k, int nb, const double *__restrict__ part, double *__restrict__ out) {
  __shared__ double red[kBS];
  for (int j = 0; j < k; j++) {
    double s = 0.0;
    for (int b = threadIdx.x; b < nb; b += blockDim.x) s += part[(long)j * nb + b];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int st = blockDim.x / 2; st > 0; st >>= 1) {
      if (threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
      __syncthreads();
    }
    if (threadIdx.x == 0) out[j] = red[0];
    __syncthreads();
  }
}

// w -= sum_j h_j V_j
__global__ void k_maxpy(long n, int k, const double *const *__restrict__ V, const double *__restrict__ h,
                        double *__restrict__ w) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    double s = w[i];
    for (int j = 0; j < k; j++) s -= h[j] * V[j][i];
    w[i] = s;
  }
}

// x += sum_j y_j Z_j
__global__ void k_mupdate(long n, int k, const double *const *__restrict__ Z, const double *__restrict__ y,
                          double *__restrict__ x) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    double s = x[i];
    for (int j = 0; j < k; j++) s += y[j] * Z[j][i];
    x[i] = s;
  }
}

// One reduction pass of the low-synchronisation Gram-Schmidt step: for the
// basis vectors V_0..V_k (chunk blockIdx.y holds V_8c..V_8c+7) the products
// <V_i, V_k> and <V_i, w>, and <w, w>; slot i, k + 1 + i and 2k + 2 of part
// (one value per block: part[slot * gridDim.x + block]).
__global__ void k_lsgs_partial(long n, int k, const double *const *__restrict__ V, const double *__restrict__ w,
                               double *__restrict__ part) {
  __shared__ double red[17][kBS / 64];
  const int i0 = 8 * blockIdx.y;
  const int ni = min(8, k + 1 - i0);
  const double *vk = V[k];
  double aa[8] = {0, 0, 0, 0, 0, 0, 0, 0}, at[8] = {0, 0, 0, 0, 0, 0, 0, 0}, ww = 0.0;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
    const double we = w[e], ve = vk[e];
    if (blockIdx.y == 0) ww += we * we;
#pragma unroll
    for (int j = 0; j < 8; j++)
      if (j < ni) {
        const double x = V[i0 + j][e];
        aa[j] += x * ve;
        at[j] += x * we;
      }
  }
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int j = 0; j < 17; j++) {
    double v = j < 8 ? aa[j] : (j < 16 ? at[j - 8] : ww);
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    if (lane == 0) red[j][wv] = v;
  }
  __syncthreads();
  if (threadIdx.x < 17) {
    const int j = threadIdx.x;
    double t = 0.0;
    for (int q = 0; q < kBS / 64; q++) t += red[j][q];
    long slot = -1;
    if (j < 8 && j < ni) slot = i0 + j;
    else if (j >= 8 && j < 16 && j - 8 < ni) slot = k + 1 + i0 + (j - 8);
    else if (j == 16 && blockIdx.y == 0) slot = 2 * k + 2;
    if (slot >= 0) part[slot * gridDim.x + blockIdx.x] = t;
  }
}

// The update pass of the same step (coefficients c: a_0..a_k-1, t_0..t_k-1,
// then t_k', 1/rho, 1/eta): V_k <- (V_k - sum a_i V_i) / rho and
// V_k+1 = (w - sum t_i V_i - t_k' V_k) / eta, reading V_0..V_k-1 once.
__global__ void k_lsgs_update(long n, int k, double *const *__restrict__ V, const double *__restrict__ c,
                              const double *__restrict__ w) {
  const double tk = c[2 * k], irho = c[2 * k + 1], ieta = c[2 * k + 2];
  double *vk = V[k], *vn = V[k + 1];
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
    double s1 = 0.0, s2 = 0.0;
    for (int i = 0; i < k; i++) {
      const double x = V[i][e];
      s1 += c[i] * x;
      s2 += c[k + i] * x;
    }
    const double v = (vk[e] - s1) * irho;
    vk[e] = v;
    vn[e] = (w[e] - s2 - tk * v) * ieta;
  }
}

// ---------------------------------------------------------------------------
// transfers between level l-1 (coarse, mxc x myc x mzc) and l (fine, doubled)
// ---------------------------------------------------------------------------
// Work-groups of a 1-D launch go round-robin to the 8 XCDs (each with its
// own L2).  Work-group b of `total` (b-th in dispatch order) takes item
// xcd_item(b): every XCD gets one contiguous run of items, so lattice planes
// that neighbouring items share are fetched into one L2, not into several.
__device__ __forceinline__ int xcd_item(int b, int total) {
  const int per = total / 8, rem = total % 8, x = b % 8, slot = b / 8;
  return x * per + (x < rem ? x : rem) + slot;
}

// yf += P xc: fine node I takes coarse node I/2 (I even) or the mean of
// coarse nodes (I-1)/2 and (I+1)/2 (I odd), tensor product, 3 components.
// Launch: bpp work-groups per fine plane, planes in z order, items mapped to
// XCDs by xcd_item (one 32-bit division per node; 32-bit indices, the host
// checks 3 N fits).
template <typename T>
__global__ void k_prolong_add(int Xc, int Yc, int Xf, int Yf, int bpp, const T *__restrict__ xc,
                              T *__restrict__ yf) {
  const int q = xcd_item(blockIdx.x, gridDim.x);
  const int K = q / bpp;
  const int t = (q - K * bpp) * blockDim.x + threadIdx.x;
  if (t >= Xf * Yf) return;
  const int J = t / Xf, I = t - J * Xf;
  int ci[2], cj[2], ck[2];
  T wi[2], wj[2], wk[2];
  int ni, nj, nk;
  if (I & 1) { ci[0] = I >> 1; ci[1] = ci[0] + 1; wi[0] = wi[1] = T(0.5); ni = 2; } else { ci[0] = I >> 1; wi[0] = T(1); ni = 1; }
  if (J & 1) { cj[0] = J >> 1; cj[1] = cj[0] + 1; wj[0] = wj[1] = T(0.5); nj = 2; } else { cj[0] = J >> 1; wj[0] = T(1); nj = 1; }
  if (K & 1) { ck[0] = K >> 1; ck[1] = ck[0] + 1; wk[0] = wk[1] = T(0.5); nk = 2; } else { ck[0] = K >> 1; wk[0] = T(1); nk = 1; }
  T s0 = 0, s1 = 0, s2 = 0;
  for (int c = 0; c < nk; c++)
    for (int b = 0; b < nj; b++)
      for (int a = 0; a < ni; a++) {
        const int g = 3 * (ci[a] + Xc * (cj[b] + Yc * ck[c]));
        const T w = wi[a] * wj[b] * wk[c];
        s0 += w * xc[g];
        s1 += w * xc[g + 1];
        s2 += w * xc[g + 2];
      }
  const int f = 3 * (t + Xf * Yf * K);
  yf[f] += s0;
  yf[f + 1] += s1;
  yf[f + 2] += s2;
}

// yc = P^T xf, gathered per coarse node from fine nodes 2i-1, 2i, 2i+1
// (weights 1/2, 1, 1/2 per direction).  Launch as k_prolong_add, over the
// coarse planes.
template <typename T>
__global__ void k_restrict(int Xc, int Yc, int Xf, int Yf, int Zf, int bpp, const T *__restrict__ xf,
                           T *__restrict__ yc) {
  const int q = xcd_item(blockIdx.x, gridDim.x);
  const int k = q / bpp;
  const int t = (q - k * bpp) * blockDim.x + threadIdx.x;
  if (t >= Xc * Yc) return;
  const int j = t / Xc, i = t - j * Xc;
  T s0 = 0, s1 = 0, s2 = 0;
  for (int dk = -1; dk <= 1; dk++) {
    const int K = 2 * k + dk;
    if (K < 0 || K >= Zf) continue;
    const T wk = dk == 0 ? T(1) : T(0.5);
    for (int dj = -1; dj <= 1; dj++) {
      const int J = 2 * j + dj;
      if (J < 0 || J >= Yf) continue;
      const T wj = dj == 0 ? T(1) : T(0.5);
      for (int di = -1; di <= 1; di++) {
        const int I = 2 * i + di;
        if (I < 0 || I >= Xf) continue;
        const T w = (di == 0 ? T(1) : T(0.5)) * wj * wk;
        const int f = 3 * (I + Xf * (J + Yf * K));
        s0 += w * xf[f];
        s1 += w * xf[f + 1];
        s2 += w * xf[f + 2];
      }
    }
  }
  const int g = 3 * (t + Xc * Yc * k);
  yc[g] = s0;
  yc[g + 1] = s1;
  yc[g + 2] = s2;
}

// Q2 embedding (transfer option 1): the fine nodal values of the coarse Q2
// function.  Per direction, fine node f = 2c coincides with coarse node c;
// an odd fine node between c and c+1 lies at xi = -1/2 (c even) or +1/2
// (c odd) of the coarse element and takes the quadratic Lagrange weights
// (3/8, 3/4, -1/8) from the element's nodes.
__device__ __forceinline__ int q2_weights(long f, long *c, double *w) {
  if (!(f & 1)) { c[0] = f >> 1; w[0] = 1.0; return 1; }
  const long cl = f >> 1;  // coarse node left of f
  if (!(cl & 1)) {         // xi = -1/2 in element [cl, cl+2]
    c[0] = cl; c[1] = cl + 1; c[2] = cl + 2;
    w[0] = 0.375; w[1] = 0.75; w[2] = -0.125;
  } else {                 // xi = +1/2 in element [cl-1, cl+1]
    c[0] = cl - 1; c[1] = cl; c[2] = cl + 1;
    w[0] = -0.125; w[1] = 0.75; w[2] = 0.375;
  }
  return 3;
}

template <typename T>
__global__ void k_prolong_add_q2(long Xc, long Yc, long Xf, long Yf, long Zf, const T *__restrict__ xc,
                                 T *__restrict__ yf) {
  const long nf = Xf * Yf * Zf;
  for (long f = blockIdx.x * (long)blockDim.x + threadIdx.x; f < nf; f += (long)gridDim.x * blockDim.x) {
    const long I = f % Xf, J = (f / Xf) % Yf, K = f / (Xf * Yf);
    long ci[3], cj[3], ck[3];
    double wi[3], wj[3], wk[3];
    const int ni = q2_weights(I, ci, wi), nj = q2_weights(J, cj, wj), nk = q2_weights(K, ck, wk);
    T s0 = 0, s1 = 0, s2 = 0;
    for (int c = 0; c < nk; c++)
      for (int b = 0; b < nj; b++)
        for (int a = 0; a < ni; a++) {
          const long g = ci[a] + Xc * (cj[b] + Yc * ck[c]);
          const T w = (T)(wi[a] * wj[b] * wk[c]);
          s0 += w * xc[3 * g];
          s1 += w * xc[3 * g + 1];
          s2 += w * xc[3 * g + 2];
        }
    yf[3 * f] += s0;
    yf[3 * f + 1] += s1;
    yf[3 * f + 2] += s2;
  }
}

// transposed Q2 embedding, gathered per coarse node: per direction the fine
// offsets from 2c with weights {-3: -1/8, -1: 3/8, 0: 1, 1: 3/8, 3: -1/8}
// (c even) or {-1: 3/4, 0: 1, 1: 3/4} (c odd)
__device__ __forceinline__ int q2_rweights(long c, int *o, double *w) {
  if (c & 1) {
    o[0] = -1; o[1] = 0; o[2] = 1;
    w[0] = 0.75; w[1] = 1.0; w[2] = 0.75;
    return 3;
  }
  o[0] = -3; o[1] = -1; o[2] = 0; o[3] = 1; o[4] = 3;
  w[0] = -0.125; w[1] = 0.375; w[2] = 1.0; w[3] = 0.375; w[4] = -0.125;
  return 5;
}

template <typename T>
__global__ void k_restrict_q2(long Xc, long Yc, long Zc, long Xf, long Yf, long Zf, const T *__restrict__ xf,
                              T *__restrict__ yc) {
  const long ncn = Xc * Yc * Zc;
  for (long g = blockIdx.x * (long)blockDim.x + threadIdx.x; g < ncn; g += (long)gridDim.x * blockDim.x) {
    const long i = g % Xc, j = (g / Xc) % Yc, k = g / (Xc * Yc);
    int oi[5], oj[5], ok[5];
    double wi[5], wj[5], wk[5];
    const int ni = q2_rweights(i, oi, wi), nj = q2_rweights(j, oj, wj), nk = q2_rweights(k, ok, wk);
    T s0 = 0, s1 = 0, s2 = 0;
    for (int c = 0; c < nk; c++) {
      const long K = 2 * k + ok[c];
      if (K < 0 || K >= Zf) continue;
      for (int b = 0; b < nj; b++) {
        const long J = 2 * j + oj[b];
        if (J < 0 || J >= Yf) continue;
        for (int a = 0; a < ni; a++) {
          const long I = 2 * i + oi[a];
          if (I < 0 || I >= Xf) continue;
          const T w = (T)(wi[a] * wj[b] * wk[c]);
          const long f = I + Xf * (J + Yf * K);
          s0 += w * xf[3 * f];
          s1 += w * xf[3 * f + 1];
          s2 += w * xf[3 * f + 2];
        }
      }
    }
    yc[3 * g] = s0;
    yc[3 * g + 1] = s1;
    yc[3 * g + 2] = s2;
  }
}

// coarse-element viscosity from the 8 x 27 quadrature values of the children
// (quadrature-point-major input ef[q * nef + e], element-major output
// ec[27 E + q]): mode 0 arithmetic, 1 geometric,
// 2 harmonic mean, one value per coarse element
__global__ void k_coarsen_eta(int mxc, int myc, int mzc, int mode, const double *__restrict__ ef,
                              double *__restrict__ ec) {
  const long nec = (long)mxc * myc * mzc, nef = 8 * nec;
  const int mxf = 2 * mxc, myf = 2 * myc;
  for (long E = blockIdx.x * (long)blockDim.x + threadIdx.x; E < nec; E += (long)gridDim.x * blockDim.x) {
    const long ei = E % mxc, ej = (E / mxc) % myc, ek = E / ((long)mxc * myc);
    double s = 0.0;
    for (int c = 0; c < 2; c++)
      for (int b = 0; b < 2; b++)
        for (int a = 0; a < 2; a++) {
          const long ef_e = (2 * ei + a) + (long)mxf * ((2 * ej + b) + (long)myf * (2 * ek + c));
          for (int q = 0; q < 27; q++) {
            const double v = ef[(long)q * nef + ef_e];
            s += mode == 0 ? v : (mode == 1 ? log(v) : 1.0 / v);
          }
        }
    s *= 1.0 / 216.0;
    const double m = mode == 0 ? s : (mode == 1 ? exp(s) : 1.0 / s);
    for (int q = 0; q < 27; q++) ec[27 * E + q] = m;
  }
}

// quadrature-point-local coarsening (modes 3 arithmetic, 4 geometric): coarse
// quadrature point (qa, qb, qc) averages the 27 fine values of each child
// element containing it (per direction: child 0 for qa = 0, child 1 for
// qa = 2, both for the midpoint qa = 1), keeping the sub-element variation
__global__ void k_coarsen_eta_qp(int mxc, int myc, int mzc, int mode, const double *__restrict__ ef,
                                 double *__restrict__ ec) {
  const long n = 27L * mxc * myc * mzc, nef = 8L * mxc * myc * mzc;
  const int mxf = 2 * mxc, myf = 2 * myc;
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
    const long E = t / 27;
    const int q = (int)(t - 27 * E);
    const int qa = q % 3, qb = (q / 3) % 3, qc = q / 9;
    const long ei = E % mxc, ej = (E / mxc) % myc, ek = E / ((long)mxc * myc);
    double s = 0.0;
    int cnt = 0;
    for (int c = (qc == 2 ? 1 : 0); c <= (qc == 0 ? 0 : 1); c++)
      for (int b = (qb == 2 ? 1 : 0); b <= (qb == 0 ? 0 : 1); b++)
        for (int a = (qa == 2 ? 1 : 0); a <= (qa == 0 ? 0 : 1); a++) {
          const long ef_e = (2 * ei + a) + (long)mxf * ((2 * ej + b) + (long)myf * (2 * ek + c));
          for (int r = 0; r < 27; r++) {
            const double v = ef[(long)r * nef + ef_e];
            s += mode == 3 ? v : log(v);
          }
          cnt += 27;
        }
    s /= (double)cnt;
    ec[t] = mode == 3 ? s : exp(s);
  }
}

// ---------------------------------------------------------------------------
// Coefficient hierarchy from quadrature-point viscosity
// (SwarmUpdateGaussPropertiesLocalL2Projection_Q1_MPntPStokes_Hierarchy,
// material_point_utils.c:1267-1420, with the fine field built from the
// quadrature points, 1138-1265): a nodal field on each level's Q2 node
// lattice, carried down by scaled restriction or injection, interpolated to
// the level's quadrature points from the 8 element corners and clamped to the
// element's nodal range (894-1000).
// ---------------------------------------------------------------------------
// Q1 factor of corner side s (0: xi = -1, 1: xi = +1) at Gauss abscissa a
__device__ __forceinline__ double q1_factor(int s, int a) {
  const double x = a == 0 ? -kS : (a == 1 ? 0.0 : kS);
  return s ? 0.5 * (1.0 + x) : 0.5 * (1.0 - x);
}

// projection weight of quadrature point q for local node (i, j, k): the Q1
// corner basis at the 8 corners, 1 at the other 19 nodes
__device__ __forceinline__ double q1_proj_weight(int i, int j, int k, int q) {
  if ((i | j | k) & 1) return 1.0;
  return q1_factor(i >> 1, q % 3) * q1_factor(j >> 1, (q / 3) % 3) * q1_factor(k >> 1, q / 9);
}

// per lattice coordinate: the (element, local index) pairs containing it,
// in increasing element order
__device__ __forceinline__ int lattice_owners(long I, int m, long *el, int *loc) {
  if (I & 1) {
    el[0] = (I - 1) / 2; loc[0] = 1;
    return 1;
  }
  int n = 0;
  if (I / 2 - 1 >= 0) { el[n] = I / 2 - 1; loc[n] = 2; n++; }
  if (I / 2 < m) { el[n] = I / 2; loc[n] = 0; n++; }
  return n;
}

// fine nodal field from qp-major viscosity eq[q * nel + e]; one thread per
// node gathers its (at most 8) elements in increasing order, each element's
// sum over its quadrature points in order (the reference's scatter order)
__global__ void k_q1_nodal(int mx, int my, int mz, const double *__restrict__ eq, double *__restrict__ nodal) {
  const long X = 2L * mx + 1, Y = 2L * my + 1, N = X * Y * (2L * mz + 1);
  const long nel = (long)mx * my * mz;
  for (long g = blockIdx.x * (long)blockDim.x + threadIdx.x; g < N; g += (long)gridDim.x * blockDim.x) {
    const long I = g % X, J = (g / X) % Y, K = g / (X * Y);
    long ex[2], ey[2], ez[2];
    int lx[2], ly[2], lz[2];
    const int nx = lattice_owners(I, mx, ex, lx), ny = lattice_owners(J, my, ey, ly),
              nz = lattice_owners(K, mz, ez, lz);
    double A = 0.0, B = 0.0;
    for (int c = 0; c < nz; c++)
      for (int b = 0; b < ny; b++)
        for (int a = 0; a < nx; a++) {
          const long e = ex[a] + mx * (ey[b] + (long)my * ez[c]);
          double Ae = 0.0, Be = 0.0;
          for (int q = 0; q < 27; q++) {
            const double w = q1_proj_weight(lx[a], ly[b], lz[c], q);
            Ae += w * eq[(long)q * nel + e];
            Be += w;
          }
          A += Ae;
          B += Be;
        }
    nodal[g] = A / B;
  }
}

// next coarser nodal field: (P^T f) * (1 / (P^T 1)) with the scalar DMDA Q1
// lattice interpolation P, or injection
__global__ void k_q1_restrict(int mxc, int myc, int mzc, int inject, const double *__restrict__ nf,
                              double *__restrict__ nc) {
  const long Xc = 2L * mxc + 1, Yc = 2L * myc + 1, Nc = Xc * Yc * (2L * mzc + 1);
  const long Xf = 4L * mxc + 1, Yf = 4L * myc + 1, Zf = 4L * mzc + 1;
  for (long g = blockIdx.x * (long)blockDim.x + threadIdx.x; g < Nc; g += (long)gridDim.x * blockDim.x) {
    const long I = g % Xc, J = (g / Xc) % Yc, K = g / (Xc * Yc);
    if (inject) {
      nc[g] = nf[2 * I + Xf * (2 * J + Yf * 2 * K)];
      continue;
    }
    double a = 0.0, s = 0.0;
    for (int dk = -1; dk <= 1; dk++) {
      const long Kf = 2 * K + dk;
      if (Kf < 0 || Kf >= Zf) continue;
      for (int dj = -1; dj <= 1; dj++) {
        const long Jf = 2 * J + dj;
        if (Jf < 0 || Jf >= Yf) continue;
        for (int di = -1; di <= 1; di++) {
          const long If = 2 * I + di;
          if (If < 0 || If >= Xf) continue;
          const double w = (di ? 0.5 : 1.0) * (dj ? 0.5 : 1.0) * (dk ? 0.5 : 1.0);
          a += w * nf[If + Xf * (Jf + Yf * Kf)];
          s += w;
        }
      }
    }
    nc[g] = a * (1.0 / s);
  }
}

// element-major quadrature values of a level from its nodal field: Q1
// interpolation from the element corners, clamped to the range of the 27
// element nodes
__global__ void k_q1_to_qp(int mx, int my, int mz, const double *__restrict__ nodal, double *__restrict__ ec) {
  const long X = 2L * mx + 1, XY = X * (2L * my + 1);
  const long n = 27L * mx * my * mz;
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
    const long e = t / 27;
    const int q = (int)(t - 27 * e);
    const long ei = e % mx, ej = (e / mx) % my, ek = e / ((long)mx * my);
    const long base = 2 * ei + X * (2 * ej) + XY * (2 * ek);
    double lo = 1.0e32, hi = -1.0e32;
    for (int k = 0; k < 3; k++)
      for (int j = 0; j < 3; j++)
        for (int i = 0; i < 3; i++) {
          const double v = nodal[base + i + X * j + XY * k];
          lo = fmin(lo, v);
          hi = fmax(hi, v);
        }
    double v = 0.0;
    for (int c = 0; c < 2; c++)
      for (int b = 0; b < 2; b++)
        for (int a = 0; a < 2; a++)
          v += q1_proj_weight(2 * a, 2 * b, 2 * c, q) * nodal[base + 2 * a + X * 2 * b + XY * 2 * c];
    ec[t] = fmin(fmax(v, lo), hi);
  }
}

// the same nodal field from material points: point p of element elem[p] at
// reference coordinates xi[3p..] with value val[p] adds its weights to the
// 27 nodes of its element (_SwarmUpdateGaussPropertiesLocalL2Projection
// Q1_MPntPStokes, material_point_utils.c:604-705); FP64 atomics, so the sums
// are formed in no fixed order
__global__ void k_q1_points_accum(long npts, int mx, int my, const long *__restrict__ elem,
                                  const double *__restrict__ xi, const double *__restrict__ val,
                                  double *__restrict__ A, double *__restrict__ B) {
  const long X = 2L * mx + 1, XY = X * (2L * my + 1);
  for (long p = blockIdx.x * (long)blockDim.x + threadIdx.x; p < npts; p += (long)gridDim.x * blockDim.x) {
    const long e = elem[p];
    const long ei = e % mx, ej = (e / mx) % my, ek = e / ((long)mx * my);
    const long base = 2 * ei + X * (2 * ej) + XY * (2 * ek);
    const double x0 = xi[3 * p], x1 = xi[3 * p + 1], x2 = xi[3 * p + 2], v = val[p];
    const double f0[2] = {0.5 * (1.0 - x0), 0.5 * (1.0 + x0)}, f1[2] = {0.5 * (1.0 - x1), 0.5 * (1.0 + x1)},
                 f2[2] = {0.5 * (1.0 - x2), 0.5 * (1.0 + x2)};
    for (int k = 0; k < 3; k++)
      for (int j = 0; j < 3; j++)
        for (int i = 0; i < 3; i++) {
          const double w = ((i | j | k) & 1) ? 1.0 : f0[i >> 1] * f1[j >> 1] * f2[k >> 1];
          const long node = base + i + X * j + XY * k;
          atomicAdd(&A[node], w * v);
          atomicAdd(&B[node], w);
        }
  }
}

__global__ void k_divide(long n, const double *__restrict__ a, const double *__restrict__ b, double *__restrict__ c) {
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x)
    c[t] = a[t] / b[t];
}

// ---------------------------------------------------------------------------
// Galerkin coarse operators stored as 125-point nodal stencils with 3x3
// blocks, offset-major for coalesced reads: S[((o * 9) + 3 r + c) * N + m] is
// the (r, c) block entry coupling node m to node m + off(o), off(o) in
// [-2, 2]^3, o = (oi+2) + 5 (oj+2) + 25 (ok+2).
// ---------------------------------------------------------------------------
// y = S x.  25 threads per node, one per (oz, oy) offset pair with the five
// x offsets in a loop, 16 nodes per work-group: consecutive threads read
// consecutive nodes of one offset row (coalesced), and there are enough
// waves to cover memory latency on the small coarse levels; the 25 partial
// sums are reduced through LDS.
// kCheb: the Chebyshev-Jacobi step is applied to the product as it is
// formed, pkp1 = os dinv (b - A x) + om1 pkm1 + om x (x = p_k), and y is not
// written (the same arithmetic as the product followed by k_cheb_step);
// pkp1 may be pkm1 (in place), pkm1 == nullptr stands for zero.
constexpr int kStNodes = 16;
template <typename T, bool kCheb = false>
__global__ void __launch_bounds__(25 * kStNodes)
    k_stencil_apply(long X, long Y, long Z, const T *__restrict__ S, const T *__restrict__ x,
                    T *__restrict__ y, const T *__restrict__ b = nullptr, const T *__restrict__ dinv = nullptr,
                    const T *pkm1 = nullptr, T os = T(0), T om1 = T(0), T om = T(0),
                    T *pkp1 = nullptr) {
  __shared__ T red[25][kStNodes][3];
  const long N = X * Y * Z;
  const int t = threadIdx.x;
  const int ml = t % kStNodes, g = t / kStNodes;
  const int ok = g / 5 - 2, oj = g % 5 - 2;
  const long m = (long)blockIdx.x * kStNodes + ml;
  T a0 = 0, a1 = 0, a2 = 0;
  if (m < N) {
    const long i = m % X, j = (m / X) % Y, k = m / (X * Y);
    const long kk = k + ok, jj = j + oj;
    if (kk >= 0 && kk < Z && jj >= 0 && jj < Y) {
#pragma unroll
      for (int oi = -2; oi <= 2; oi++) {
        const long ii = i + oi;
        if (ii < 0 || ii >= X) continue;
        const int o = (oi + 2) + 5 * (oj + 2) + 25 * (ok + 2);
        const long n = ii + X * (jj + Y * kk);
        const T x0 = x[3 * n], x1 = x[3 * n + 1], x2 = x[3 * n + 2];
        const T *b = S + (long)o * 9 * N + m;
        a0 += b[0] * x0 + b[N] * x1 + b[2 * N] * x2;
        a1 += b[3 * N] * x0 + b[4 * N] * x1 + b[5 * N] * x2;
        a2 += b[6 * N] * x0 + b[7 * N] * x1 + b[8 * N] * x2;
      }
    }
  }
  red[g][ml][0] = a0;
  red[g][ml][1] = a1;
  red[g][ml][2] = a2;
  __syncthreads();
  if (g < 3 && m < N) {
    T acc = 0;
    for (int q = 0; q < 25; q++) acc += red[q][ml][g];
    const long i = 3 * m + g;
    if constexpr (kCheb) {
      const T z = (b[i] - acc) * dinv[i];
      const T prev = pkm1 ? pkm1[i] : T(0);
      pkp1[i] = os * z + om1 * prev + om * x[i];
    } else {
      y[i] = acc;
    }
  }
}

// probe vector: v[3n + d] = 1 for the nodes of colour (ci, cj, ck) (indices
// modulo 5), 0 elsewhere
__global__ void k_probe(long X, long Y, long Z, int ci, int cj, int ck, int d, double *__restrict__ v) {
  const long N = X * Y * Z;
  for (long n = blockIdx.x * (long)blockDim.x + threadIdx.x; n < N; n += (long)gridDim.x * blockDim.x) {
    const long i = n % X, j = (n / X) % Y, k = n / (X * Y);
    const bool on = (i % 5 == ci) && (j % 5 == cj) && (k % 5 == ck);
    v[3 * n] = (on && d == 0) ? 1.0 : 0.0;
    v[3 * n + 1] = (on && d == 1) ? 1.0 : 0.0;
    v[3 * n + 2] = (on && d == 2) ? 1.0 : 0.0;
  }
}

// row m of the probed product w = A v holds column (n, d) of A, n the unique
// node of the colour within +-2 of m
__global__ void k_probe_scatter(long X, long Y, long Z, int ci, int cj, int ck, int d,
                                const double *__restrict__ w, double *__restrict__ S) {
  const long N = X * Y * Z;
  for (long m = blockIdx.x * (long)blockDim.x + threadIdx.x; m < N; m += (long)gridDim.x * blockDim.x) {
    const long i = m % X, j = (m / X) % Y, k = m / (X * Y);
    int oi = (int)(((ci - i) % 5 + 5) % 5), oj = (int)(((cj - j) % 5 + 5) % 5), ok = (int)(((ck - k) % 5 + 5) % 5);
    if (oi > 2) oi -= 5;
    if (oj > 2) oj -= 5;
    if (ok > 2) ok -= 5;
    if (i + oi < 0 || i + oi >= X || j + oj < 0 || j + oj >= Y || k + ok < 0 || k + ok >= Z) continue;
    const int o = (oi + 2) + 5 * (oj + 2) + 25 * (ok + 2);
    for (int r = 0; r < 3; r++) S[((long)o * 9 + 3 * r + d) * N + m] = w[3 * m + r];
  }
}

// the same probe scattered into a dense column-major matrix of order 3N:
// entry (3m + r, 3n + d) = w[3m + r], n the colour's node within +-2 of m
__global__ void k_probe_scatter_dense(long X, long Y, long Z, int ci, int cj, int ck, int d,
                                      const double *__restrict__ w, double *__restrict__ A) {
  const long N = X * Y * Z, n3 = 3 * N;
  for (long m = blockIdx.x * (long)blockDim.x + threadIdx.x; m < N; m += (long)gridDim.x * blockDim.x) {
    const long i = m % X, j = (m / X) % Y, k = m / (X * Y);
    int oi = (int)(((ci - i) % 5 + 5) % 5), oj = (int)(((cj - j) % 5 + 5) % 5), ok = (int)(((ck - k) % 5 + 5) % 5);
    if (oi > 2) oi -= 5;
    if (oj > 2) oj -= 5;
    if (ok > 2) ok -= 5;
    if (i + oi < 0 || i + oi >= X || j + oj < 0 || j + oj >= Y || k + ok < 0 || k + ok >= Z) continue;
    const long col = 3 * (m + oi + X * (oj + Y * ok)) + d;
    for (int r = 0; r < 3; r++) A[(3 * m + r) + col * n3] = w[3 * m + r];
  }
}

// Dirichlet dof g = 3m + d of a Galerkin level: identity row (its row and
// column entries are already zero from the masked probes)
__global__ void k_stencil_bc_identity(long N, const long *__restrict__ idx, long nbc, double *__restrict__ S) {
  const int oc = 62;  // offset (0,0,0)
  for (long k = blockIdx.x * (long)blockDim.x + threadIdx.x; k < nbc; k += (long)gridDim.x * blockDim.x) {
    const long g = idx[k], m = g / 3;
    const int d = (int)(g - 3 * m);
    S[((long)oc * 9 + 4 * d) * N + m] = 1.0;
  }
}

// l1 row sums of a Galerkin stencil: sum over offsets and columns of |S|
__global__ void k_stencil_l1(long N, const double *__restrict__ S, double *__restrict__ dg) {
  for (long m = blockIdx.x * (long)blockDim.x + threadIdx.x; m < N; m += (long)gridDim.x * blockDim.x)
    for (int r = 0; r < 3; r++) {
      double s = 0.0;
      for (int o = 0; o < 125; o++)
        for (int c = 0; c < 3; c++) s += fabs(S[((long)o * 9 + 3 * r + c) * N + m]);
      dg[3 * m + r] = s;
    }
}

__global__ void k_stencil_diag(long N, const double *__restrict__ S, double *__restrict__ dg) {
  const int oc = 62;  // offset (0,0,0)
  for (long m = blockIdx.x * (long)blockDim.x + threadIdx.x; m < N; m += (long)gridDim.x * blockDim.x)
    for (int r = 0; r < 3; r++) dg[3 * m + r] = S[((long)oc * 9 + 4 * r) * N + m];
}

}  // namespace

// ---------------------------------------------------------------------------
// host-side multigrid object
// ---------------------------------------------------------------------------
struct MGLevel {
  int mx = 0, my = 0, mz = 0;
  ptmf_op op = nullptr;
  long n = 0;  // 3 * nnodes
  double *diag = nullptr, *dinv64 = nullptr;
  float *dinv32 = nullptr;
  double emin = 0, emax = 0;
  bool have_eta = false, have_bc = false;
  bool galerkin = false;  // level operator = R A_{l+1} P, stored as a stencil
  double *S64 = nullptr;
  float *S32 = nullptr;
  // work vectors
  double *b64 = nullptr, *x64 = nullptr, *r64 = nullptr, *p64[3] = {nullptr, nullptr, nullptr};
  float *b32 = nullptr, *x32 = nullptr, *r32 = nullptr, *p32[3] = {nullptr, nullptr, nullptr};
};

}  // namespace ptmf

namespace ptmf {
// FGMRES basis: V (restart+1 vectors), Z (restart preconditioned vectors),
// a work vector, and device arrays of the vector pointers for the fused
// multi-vector kernels.
struct KrylovSpace {
  int restart = 0;
  long n = 0;
  std::vector<double *> V, Z;
  double *w = nullptr;
  const double **dV = nullptr, **dZ = nullptr;
  void release() {
    for (double *p : V) (void)hipFree(p);
    for (double *p : Z) (void)hipFree(p);
    (void)hipFree(w);
    (void)hipFree((void *)dV);
    (void)hipFree((void *)dZ);
    V.clear();
    Z.clear();
    w = nullptr;
    dV = dZ = nullptr;
    restart = 0;
    n = 0;
  }
};
}  // namespace ptmf

struct ptmf_mg_s {
  int L = 0;
  ptmf::MGLevel lev[16];
  ptmf_mg_options opt;
  bool is_setup = false;
  hipStream_t stream = nullptr;
  // reductions
  double *red_part = nullptr, *red_out = nullptr, *red_host = nullptr;
  double *coef_dev = nullptr, *coef_host = nullptr;  // Krylov coefficients (pinned upload)
  // dense coarse factor
  rocblas_handle blas = nullptr;
  ptmf::MFFactor *mf = nullptr; // multifrontal Cholesky of level 0 (sparse coarse factor)
  double *Adense = nullptr;     // explicit inverse, full symmetric (column-major)
  double *dense_part = nullptr; // partial sums of the inverse's product
  int ndense = 0;
  int *dinfo = nullptr;
  // Krylov workspaces (finest level, FP64): velocity block and saddle point
  ptmf::KrylovSpace ku, ks;
  double *pcg[4] = {nullptr, nullptr, nullptr, nullptr};
  // saddle point: inverse Schur blocks (16 per element) and a velocity temporary
  double *sMinv = nullptr, *stmp = nullptr;
  double *sx2 = nullptr, *sy2 = nullptr;  // [u | p] temporaries of the preconditioner's A12 product
  float *sx2f = nullptr, *sy2f = nullptr;  // the same in FP32 (FP32 velocity V-cycle)
  double *kvb = nullptr, *kvx = nullptr;   // fixed V-cycle input/output of the FGMRES preconditioners
  bool stokes_ready = false;
  // captured V-cycles (levels_mg == L), one per (b, x) pair of device
  // pointers the cycle was called with: a replay reads b and writes x
  // directly, no staging copies
  static constexpr int kGraphs = 8;
  struct VGraph {
    const double *b = nullptr;
    double *x = nullptr;
    hipGraph_t g = nullptr;
    hipGraphExec_t ex = nullptr;
    unsigned long used = 0;
  } vg[kGraphs];
  unsigned long vtick = 0;
  int vcalls = 0;
  // what each level's MatMult nodes were captured with: a later change of a
  // level operator's coefficients, tiling, kernel variant or Dirichlet list
  // (through ptmf_mg_level_op) makes the graph stale, and it is dropped
  struct GraphKey {
    unsigned long state = 0;
    int TX = 0, TY = 0, ZC = 0, variant = 0, wg = 1;
    const void *faces = nullptr, *bcidx = nullptr, *et64 = nullptr, *et32 = nullptr;
    long nbc = 0;
  } gkey[16];
  int s_inner_its = 1;       // A11^-1 in the preconditioner: 1 = one V-cycle, k > 1 = up to k
  double s_inner_rtol = 0.0; // flexible-CG iterations to relative tolerance s_inner_rtol
  double *fcg[4] = {nullptr, nullptr, nullptr, nullptr};  // flexible CG workspace: r, z, p, q
  std::vector<double> history;
};

namespace ptmf {
namespace {

int dalloc(void **p, size_t bytes) { return ptmf_malloc(p, bytes); }

template <typename T>
int axpby(hipStream_t s, long n, T a, const T *x, T b, T *y) {
  k_axpby<T><<<nblocks(n), kBS, 0, s>>>(n, a, x, b, y);
  PTMF_HIP(hipGetLastError());
  return PTMF_OK;
}

template <typename Ts, typename Td>
int cast(hipStream_t s, long n, const Ts *x, Td *y) {
  k_cast<Ts, Td><<<nblocks(n), kBS, 0, s>>>(n, x, y);
  PTMF_HIP(hipGetLastError());
  return PTMF_OK;
}

// k dot products <V_j, w> (dV: device array of k vector pointers); results
// on the host in out[0..k) and on the device in mg->red_out[0..k)
int mdot(ptmf_mg_s *mg, long n, int k, const double *const *dV, const double *w, double *out) {
  PTMF_CHECK(k > 0 && k <= kMaxDots, PTMF_ERR_ARG, "mdot: %d vectors", k);
  for (int j0 = 0; j0 < k; j0 += 8) {
    const int nk = k - j0 < 8 ? k - j0 : 8;
    k_mdot_partial<<<kRedBlocks, kBS, 0, mg->stream>>>(n, j0, nk, dV, w, mg->red_part);
    PTMF_HIP(hipGetLastError());
  }
  k_mdot_final<<<1, kBS, 0, mg->stream>>>(k, kRedBlocks, mg->red_part, mg->red_out);
  PTMF_HIP(hipGetLastError());
  PTMF_HIP(hipMemcpyAsync(mg->red_host, mg->red_out, sizeof(double) * k, hipMemcpyDeviceToHost,
                          mg->stream));
  PTMF_HIP(hipStreamSynchronize(mg->stream));
  memcpy(out, mg->red_host, sizeof(double) * k);
  return PTMF_OK;
}

int dot(ptmf_mg_s *mg, long n, const double *a, const double *b, double *out) {
  k_dot_partial<<<kRedBlocks, kBS, 0, mg->stream>>>(n, a, b, mg->red_part);
  PTMF_HIP(hipGetLastError());
  k_mdot_final<<<1, kBS, 0, mg->stream>>>(1, kRedBlocks, mg->red_part, mg->red_out);
  PTMF_HIP(hipGetLastError());
  PTMF_HIP(hipMemcpyAsync(mg->red_host, mg->red_out, sizeof(double), hipMemcpyDeviceToHost, mg->stream));
  PTMF_HIP(hipStreamSynchronize(mg->stream));
  *out = mg->red_host[0];
  return PTMF_OK;
}

// (r.r, z.r, z.q) in one pass and one host synchronisation
int dot3(ptmf_mg_s *mg, long n, const double *r, const double *z, const double *q, double *out) {
  k_dot3_partial<<<kRedBlocks, kBS, 0, mg->stream>>>(n, r, z, q, mg->red_part);
  PTMF_HIP(hipGetLastError());
  k_mdot_final<<<1, kBS, 0, mg->stream>>>(3, kRedBlocks, mg->red_part, mg->red_out);
  PTMF_HIP(hipGetLastError());
  PTMF_HIP(hipMemcpyAsync(mg->red_host, mg->red_out, sizeof(double) * 3, hipMemcpyDeviceToHost, mg->stream));
  PTMF_HIP(hipStreamSynchronize(mg->stream));
  memcpy(out, mg->red_host, sizeof(double) * 3);
  return PTMF_OK;
}

template <typename T>
int scale(hipStream_t s, long n, T a, const T *x, T *y) {
  k_scale<T><<<nblocks(n), kBS, 0, s>>>(n, a, x, y);
  PTMF_HIP(hipGetLastError());
  return PTMF_OK;
}

template <typename T>
int matmult(const MGLevel &L, const T *x, T *y) {
  if (!L.galerkin) return launch_a11_mult<T>(L.op, x, y, true);
  const Box &b = L.op->box;
  const T *S = sizeof(T) == 8 ? (const T *)L.S64 : (const T *)L.S32;
  const long nb = (b.nnodes + kStNodes - 1) / kStNodes;
  k_stencil_apply<T><<<dim3((unsigned)nb), dim3(25 * kStNodes), 0, L.op->stream>>>(b.X, b.Y, b.Z, S, x, y);
  PTMF_HIP(hipGetLastError());
  return PTMF_OK;
}

// transfers between level l (fine) and l-1: DMDA Q1 lattice interpolation
// (transfer 0, as the reference's PCMG) or the Q2 embedding (transfer 1)
template <typename T>
int prolong_add(ptmf_mg_s *mg, int l, const T *xc, T *yf) {
  const Box &bf = mg->lev[l].op->box, &bc = mg->lev[l - 1].op->box;
  if (mg->opt.transfer == 1) {
    k_prolong_add_q2<T><<<nblocks(bf.nnodes), kBS, 0, mg->stream>>>(bc.X, bc.Y, bf.X, bf.Y, bf.Z, xc, yf);
  } else {
    PTMF_CHECK(3 * bf.nnodes < (1L << 31), PTMF_ERR_UNSUPPORTED, "prolong: level too large for 32-bit indices");
    const int bpp = (int)((bf.X * bf.Y + kBS - 1) / kBS);
    k_prolong_add<T><<<dim3((unsigned)(bpp * bf.Z)), kBS, 0, mg->stream>>>((int)bc.X, (int)bc.Y, (int)bf.X,
                                                                          (int)bf.Y, bpp, xc, yf);
  }
  PTMF_HIP(hipGetLastError());
  return PTMF_OK;
}

template <typename T>
int restrict_(ptmf_mg_s *mg, int l, const T *xf, T *yc) {
  const Box &bf = mg->lev[l].op->box, &bc = mg->lev[l - 1].op->box;
  if (mg->opt.transfer == 1) {
    k_restrict_q2<T><<<nblocks(bc.nnodes), kBS, 0, mg->stream>>>(bc.X, bc.Y, bc.Z, bf.X, bf.Y, bf.Z, xf, yc);
  } else {
    PTMF_CHECK(3 * bf.nnodes < (1L << 31), PTMF_ERR_UNSUPPORTED, "restrict: level too large for 32-bit indices");
    const int bpp = (int)((bc.X * bc.Y + kBS - 1) / kBS);
    k_restrict<T><<<dim3((unsigned)(bpp * bc.Z)), kBS, 0, mg->stream>>>((int)bc.X, (int)bc.Y, (int)bf.X, (int)bf.Y,
                                                                       (int)bf.Z, bpp, xf, yc);
  }
  PTMF_HIP(hipGetLastError());
  return PTMF_OK;
}

// Chebyshev-Jacobi smoothing on level L (precision T): x <- cheb(b, x).
// b64 (zero initial guess only): the right-hand side arrives in FP64 and the
// first step also writes its rounding into b, which the later steps read.
// x64: the result goes to x64 in FP64 (the last step stores it) and x is not
// written.  Both fold the precision casts of an FP32 V-cycle's finest level
// into the smoother's first and last kernels.
template <typename T>
int cheb(ptmf_mg_s *mg, MGLevel &L, const T *b, T *x, int its, bool nonzero, const double *b64 = nullptr,
         double *x64 = nullptr) {
  hipStream_t s = mg->stream;
  const long n = L.n;
  T *p[3], *r;
  const T *dinv;
  if (sizeof(T) == 8) {
    p[0] = (T *)L.p64[0]; p[1] = (T *)L.p64[1]; p[2] = (T *)L.p64[2];
    r = (T *)L.r64; dinv = (const T *)L.dinv64;
  } else {
    p[0] = (T *)L.p32[0]; p[1] = (T *)L.p32[1]; p[2] = (T *)L.p32[2];
    r = (T *)L.r32; dinv = (const T *)L.dinv32;
  }
  const double emin = L.emin, emax = L.emax;
  const double scale = 2.0 / (emax + emin);
  const double alpha = 1.0 - scale * emin;
  const double mu = 1.0 / alpha, omegaprod = 2.0 / alpha;
  double c[3];
  // Two iterates are live at a time: each step overwrites the iterate before
  // last in place (kernels take pkp1 == pkm1), so the recurrence needs one
  // buffer besides x.  A zero initial guess is no buffer at all (nullptr),
  // and the steps alternate between x and p[0], so with an odd number of
  // steps the last one lands in x: no fill, no copies (kernels only: the
  // V-cycle may be captured into a graph).
  T *prev = nullptr, *cur = p[0];
  PTMF_CHECK(!b64 || !nonzero, PTMF_ERR_ARG, "cheb: an FP64 right-hand side needs a zero initial guess");
  if (nonzero) {
    PTMF_TRY(matmult<T>(L, x, r));
    k_cheb_first<T><<<nblocks(n), kBS, 0, s>>>(n, b, r, x, dinv, (T)scale, cur);
    prev = x;
  } else if (b64) {
    k_cheb_first<T, double><<<nblocks(n), kBS, 0, s>>>(n, b64, nullptr, nullptr, dinv, (T)scale, cur, (T *)b);
  } else {
    k_cheb_first<T><<<nblocks(n), kBS, 0, s>>>(n, b, nullptr, nullptr, dinv, (T)scale, cur);
  }
  PTMF_HIP(hipGetLastError());
  c[0] = 1.0;  // c[km1]
  c[1] = mu;   // c[k]
  for (int i = 0; i < its; i++) {
    c[2] = 2.0 * mu * c[1] - c[0];
    const double omega = omegaprod * c[1] / c[2];
    T *out = prev ? prev : x;
    if (L.galerkin) {
      // stencil product and Chebyshev update in one launch
      const Box &bx = L.op->box;
      const T *S = sizeof(T) == 8 ? (const T *)L.S64 : (const T *)L.S32;
      const long nb = (bx.nnodes + kStNodes - 1) / kStNodes;
      k_stencil_apply<T, true><<<dim3((unsigned)nb), dim3(25 * kStNodes), 0, L.op->stream>>>(
          bx.X, bx.Y, bx.Z, S, cur, r, b, dinv, prev, (T)(omega * scale), (T)(1.0 - omega), (T)omega, out);
    } else {
      // MatMult without its Dirichlet-copy launch; the update takes y = x there
      PTMF_TRY(launch_a11_mult<T>(L.op, cur, r, 2));
      if (x64 && i == its - 1) {
        // the last step stores the result in FP64
        k_cheb_step<T, double><<<nblocks(n), kBS, 0, s>>>(n, b, r, dinv, prev, cur, (T)(omega * scale),
                                                           (T)(1.0 - omega), (T)omega, x64, L.op->bcnode);
        PTMF_HIP(hipGetLastError());
        return PTMF_OK;
      }
      k_cheb_step<T><<<nblocks(n), kBS, 0, s>>>(n, b, r, dinv, prev, cur, (T)(omega * scale), (T)(1.0 - omega),
                                                 (T)omega, out, L.op->bcnode);
    }
    PTMF_HIP(hipGetLastError());
    c[0] = c[1];
    c[1] = c[2];
    prev = cur;
    cur = out;
  }
  if (x64) return cast<T, double>(s, n, cur, x64);
  if (cur != x) {
    k_copy<T><<<nblocks(n), kBS, 0, s>>>(n, cur, x);
    PTMF_HIP(hipGetLastError());
  }
  return PTMF_OK;
}

// dense Cholesky solve on level 0: x = A0^-1 b (FP64)
// x = A0^-1 b with the explicit inverse (lower triangle, formed once from the
// Cholesky factor): one symmetric matrix-vector product instead of two
// latency-bound triangular solves per V-cycle
int dense_solve(ptmf_mg_s *mg, const double *b, double *x) {
  const long n = mg->ndense;
  const int nchunks = (int)((n + kGvCols - 1) / kGvCols);
  const dim3 grid((unsigned)((n + kGvRows - 1) / kGvRows), (unsigned)nchunks);
  k_gemv_part<<<grid, kGvRows, 0, mg->stream>>>(n, n, mg->Adense, n, b, mg->dense_part);
  k_gemv_sum<<<nblocks(n), kBS, 0, mg->stream>>>(n, nchunks, mg->dense_part, x);
  PTMF_HIP(hipGetLastError());
  return PTMF_OK;
}

int vcycle64(ptmf_mg_s *mg, int l, int lc, const double *b, double *x);
int coarse_solve(ptmf_mg_s *mg, int lc, const double *b, double *x);

// V-cycle in precision T from level l down to the coarse-solve level lc;
// x is overwritten (zero initial guess), b is the right-hand side.  b64 /
// x64 (level l only): the right-hand side arrives and the result leaves in
// FP64, with b and x the level's buffers in T (the casts are folded into the
// smoother's first and last kernels).
template <typename T>
int vcycle(ptmf_mg_s *mg, int l, int lc, const T *b, T *x, const double *b64 = nullptr, double *x64 = nullptr) {
  MGLevel &L = mg->lev[l];
  hipStream_t s = mg->stream;
  if (l == lc) {
    if (sizeof(T) == 8) return coarse_solve(mg, lc, (const double *)b, (double *)x);
    // FP64 coarse solve of an FP32 V-cycle
    PTMF_TRY((cast<float, double>(s, L.n, (const float *)b, L.b64)));
    PTMF_TRY(coarse_solve(mg, lc, L.b64, L.x64));
    return cast<double, float>(s, L.n, L.x64, (float *)x);
  }
  MGLevel &C = mg->lev[l - 1];
  T *r = sizeof(T) == 8 ? (T *)L.r64 : (T *)L.r32;
  T *bc = sizeof(T) == 8 ? (T *)C.b64 : (T *)C.b32;
  T *xc = sizeof(T) == 8 ? (T *)C.x64 : (T *)C.x32;
  // pre-smooth from zero
  PTMF_TRY(cheb<T>(mg, L, b, x, mg->opt.pre_its, false, b64));
  // residual and restriction
  PTMF_TRY(matmult<T>(L, x, r));
  k_residual<T><<<nblocks(L.n), kBS, 0, s>>>(L.n, b, r, r);
  PTMF_HIP(hipGetLastError());
  PTMF_TRY(restrict_<T>(mg, l, r, bc));
  // coarse correction vanishing on the Dirichlet dofs: the coarse Dirichlet
  // rows are identities, so a nonzero restricted residual there would act as
  // boundary data of the correction
  if (mg->opt.bc_transfer) PTMF_TRY(launch_bc_set<T>(C.op->bcidx, C.op->nbc, (T)0, bc, s));
  PTMF_TRY(vcycle<T>(mg, l - 1, lc, bc, xc));
  PTMF_TRY(prolong_add<T>(mg, l, xc, x));
  // post-smooth from the corrected iterate
  PTMF_TRY(cheb<T>(mg, L, b, x, mg->opt.post_its, true, nullptr, x64));
  return PTMF_OK;
}

int vcycle64(ptmf_mg_s *mg, int l, int lc, const double *b, double *x) {
  return vcycle<double>(mg, l, lc, b, x);
}

// Coarse solve on level lc: dense Cholesky on level 0, otherwise FP64 PCG
// preconditioned by an FP64 V-cycle over levels lc..0.
int coarse_solve(ptmf_mg_s *mg, int lc, const double *b, double *x) {
  if (lc == 0) return mg->mf ? mf_solve(mg->mf, b, x, mg->stream) : dense_solve(mg, b, x);
  MGLevel &L = mg->lev[lc];
  hipStream_t s = mg->stream;
  const long n = L.n;
  double *r = mg->pcg[0], *z = mg->pcg[1], *p = mg->pcg[2], *q = mg->pcg[3];
  PTMF_HIP(hipMemsetAsync(x, 0, sizeof(double) * n, s));
  PTMF_HIP(hipMemcpyAsync(r, b, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
  double rr0;
  PTMF_TRY(dot(mg, n, r, r, &rr0));
  if (rr0 == 0.0) return PTMF_OK;
  // the inner V-cycle needs level lc's b/x buffers: use z as output
  PTMF_TRY(vcycle64(mg, lc, 0, r, z));
  PTMF_HIP(hipMemcpyAsync(p, z, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
  double rz;
  PTMF_TRY(dot(mg, n, r, z, &rz));
  for (int it = 0; it < mg->opt.coarse_maxit; it++) {
    PTMF_TRY(matmult<double>(L, p, q));
    double pq;
    PTMF_TRY(dot(mg, n, p, q, &pq));
    const double a = rz / pq;
    PTMF_TRY(axpby<double>(s, n, a, p, 1.0, x));
    PTMF_TRY(axpby<double>(s, n, -a, q, 1.0, r));
    double rr;
    PTMF_TRY(dot(mg, n, r, r, &rr));
    PTMF_CHECK(isfinite(rr), PTMF_ERR_DIVERGED, "coarse PCG: non-finite residual");
    if (sqrt(rr / rr0) < mg->opt.coarse_rtol) break;
    PTMF_TRY(vcycle64(mg, lc, 0, r, z));
    double rz1;
    PTMF_TRY(dot(mg, n, r, z, &rz1));
    const double beta = rz1 / rz;
    rz = rz1;
    PTMF_TRY(axpby<double>(s, n, 1.0, z, beta, p));
  }
  return PTMF_OK;
}

// largest eigenvalue of a symmetric tridiagonal matrix (bisection, Sturm count)
double tridiag_max_eig(const std::vector<double> &d, const std::vector<double> &e) {
  const int m = (int)d.size();
  double lo = 1e300, hi = -1e300;
  for (int i = 0; i < m; i++) {
    const double r = (i > 0 ? fabs(e[i - 1]) : 0.0) + (i + 1 < m ? fabs(e[i]) : 0.0);
    lo = fmin(lo, d[i] - r);
    hi = fmax(hi, d[i] + r);
  }
  for (int it = 0; it < 200; it++) {
    const double mid = 0.5 * (lo + hi);
    int cnt = 0;  // eigenvalues < mid
    double q = d[0] - mid;
    if (q < 0) cnt++;
    for (int i = 1; i < m; i++) {
      if (q == 0.0) q = 1e-300;
      q = d[i] - mid - e[i - 1] * e[i - 1] / q;
      if (q < 0) cnt++;
    }
    if (cnt == m) hi = mid; else lo = mid;
  }
  return hi;
}

// Estimate the extreme eigenvalue of D^-1 A on level l with `its` steps of
// Jacobi-preconditioned CG (Lanczos tridiagonal of the CG coefficients).
int estimate_emax(ptmf_mg_s *mg, int l, int its, double *emax) {
  MGLevel &L = mg->lev[l];
  hipStream_t s = mg->stream;
  const long n = L.n;
  double *r = L.p64[0], *z = L.p64[1], *p = L.p64[2], *q = L.r64, *x = L.x64;
  // seeded right-hand side on the host
  std::vector<double> h(n);
  unsigned long st = 0x9E3779B97F4A7C15UL ^ (unsigned long)n;
  for (long i = 0; i < n; i++) {
    st = st * 6364136223846793005UL + 1442695040888963407UL;
    h[i] = ((double)(st >> 11) / (double)(1UL << 53)) - 0.5;
  }
  PTMF_HIP(hipMemcpyAsync(r, h.data(), sizeof(double) * n, hipMemcpyHostToDevice, s));
  // l1 scaling: estimate on the interior block (the Dirichlet block, identity
  // rows over masked columns, is decoupled and gets its own scaling)
  if (mg->opt.smoother_diag == 1) PTMF_TRY(launch_bc_set<double>(L.op->bcidx, L.op->nbc, 0.0, r, s));
  PTMF_HIP(hipMemsetAsync(x, 0, sizeof(double) * n, s));
  std::vector<double> alpha, beta;
  // z = D^-1 r
  auto jac = [&](const double *a, double *o) {
    k_cheb_first<double><<<nblocks(n), kBS, 0, s>>>(n, a, nullptr, nullptr, L.dinv64, 1.0, o);
  };
  jac(r, z);
  PTMF_HIP(hipMemcpyAsync(p, z, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
  double rz;
  PTMF_TRY(dot(mg, n, r, z, &rz));
  for (int it = 0; it < its; it++) {
    PTMF_TRY(matmult<double>(L, p, q));
    double pq;
    PTMF_TRY(dot(mg, n, p, q, &pq));
    if (!(pq > 0.0)) break;
    const double a = rz / pq;
    alpha.push_back(a);
    PTMF_TRY(axpby<double>(s, n, -a, q, 1.0, r));
    jac(r, z);
    double rz1;
    PTMF_TRY(dot(mg, n, r, z, &rz1));
    const double b = rz1 / rz;
    beta.push_back(b);
    rz = rz1;
    if (!(rz > 0.0)) break;
    PTMF_TRY(axpby<double>(s, n, 1.0, z, b, p));
  }
  const int m = (int)alpha.size();
  PTMF_CHECK(m > 0, PTMF_ERR_DIVERGED, "eigen estimate failed on level %d", l);
  std::vector<double> d(m), e(m > 1 ? m - 1 : 1, 0.0);
  for (int j = 0; j < m; j++) {
    d[j] = 1.0 / alpha[j] + (j > 0 ? beta[j - 1] / alpha[j - 1] : 0.0);
    if (j + 1 < m) e[j] = sqrt(beta[j]) / alpha[j];
  }
  *emax = tridiag_max_eig(d, e);
  return PTMF_OK;
}

// With l1 scaling the decoupled Dirichlet dofs (identity rows) would sit at
// eigenvalue 1 of D^-1 A, outside the interior spectrum: scale them to the
// middle of the Chebyshev interval instead.  Point Jacobi keeps unit
// Dirichlet diagonals, as PCJACOBI does.
int place_dirichlet(ptmf_mg_s *mg, MGLevel &L) {
  if (mg->opt.smoother_diag != 1) return PTMF_OK;
  const double mid = 0.5 * (L.emin + L.emax);
  PTMF_TRY(launch_bc_set<double>(L.op->bcidx, L.op->nbc, mid, L.dinv64, mg->stream));
  PTMF_TRY(launch_bc_set<float>(L.op->bcidx, L.op->nbc, (float)mid, L.dinv32, mg->stream));
  return PTMF_OK;
}

// Galerkin operator of level l: S_l = R A_{l+1} P by probing with the 375
// colourings (node indices modulo 5, one component each); the coupling range
// of R A P stays within +-2 nodes, so every probe yields one stencil entry
// per row and component.  P acts on the coarse non-Dirichlet dofs only (the
// probes are masked and the product's Dirichlet rows dropped); Dirichlet dofs
// get identity rows, as the rediscretised operators have.
int assemble_galerkin(ptmf_mg_s *mg, int l) {
  MGLevel &Cl = mg->lev[l], &Fl = mg->lev[l + 1];
  hipStream_t s = mg->stream;
  const Box &bc = Cl.op->box;
  const long N = bc.nnodes;
  if (!Cl.S64) PTMF_TRY(dalloc((void **)&Cl.S64, sizeof(double) * 1125 * N));
  if (!Cl.S32) PTMF_TRY(dalloc((void **)&Cl.S32, sizeof(float) * 1125 * N));
  PTMF_HIP(hipMemsetAsync(Cl.S64, 0, sizeof(double) * 1125 * N, s));
  double *v = Cl.p64[0], *w = Cl.r64, *pf = Fl.p64[0], *apf = Fl.p64[1];
  for (int ck = 0; ck < 5; ck++)
    for (int cj = 0; cj < 5; cj++)
      for (int ci = 0; ci < 5; ci++)
        for (int d = 0; d < 3; d++) {
          k_probe<<<nblocks(N), kBS, 0, s>>>(bc.X, bc.Y, bc.Z, ci, cj, ck, d, v);
          PTMF_TRY(launch_bc_set<double>(Cl.op->bcidx, Cl.op->nbc, 0.0, v, s));
          PTMF_HIP(hipMemsetAsync(pf, 0, sizeof(double) * Fl.n, s));
          PTMF_TRY(prolong_add<double>(mg, l + 1, v, pf));
          PTMF_TRY(matmult<double>(Fl, pf, apf));
          PTMF_TRY(restrict_<double>(mg, l + 1, apf, w));
          PTMF_TRY(launch_bc_set<double>(Cl.op->bcidx, Cl.op->nbc, 0.0, w, s));
          k_probe_scatter<<<nblocks(N), kBS, 0, s>>>(bc.X, bc.Y, bc.Z, ci, cj, ck, d, w, Cl.S64);
          PTMF_HIP(hipGetLastError());
        }
  if (Cl.op->nbc > 0)
    k_stencil_bc_identity<<<nblocks(Cl.op->nbc), kBS, 0, s>>>(N, Cl.op->bcidx, Cl.op->nbc, Cl.S64);
  k_cast<double, float><<<nblocks(1125 * N), kBS, 0, s>>>(1125 * N, Cl.S64, Cl.S32);
  PTMF_HIP(hipGetLastError());
  Cl.galerkin = true;
  return PTMF_OK;
}

int assemble_dense(ptmf_mg_s *mg) {
  MGLevel &L = mg->lev[0];
  const long n = L.n;
  PTMF_CHECK(n <= kDenseMaxDofs, PTMF_ERR_UNSUPPORTED,
             "coarsest level has %ld dofs; the dense coarse solve takes at most %ld (coarse_factor 2: "
             "sparse multifrontal Cholesky)", n, kDenseMaxDofs);
  mg->ndense = (int)n;
  if (!mg->Adense) PTMF_TRY(dalloc((void **)&mg->Adense, sizeof(double) * n * n));
  hipStream_t s = mg->stream;
  // 375 coloured probes instead of n unit vectors: nodes of one colour
  // (i, j, k mod 5 fixed) are 5 apart, so their columns of the level
  // operator (couplings within +-2 nodes) do not overlap
  const Box &b0 = L.op->box;
  double *v = L.p64[0], *w = L.r64;
  PTMF_HIP(hipMemsetAsync(mg->Adense, 0, sizeof(double) * n * n, s));
  for (int ck = 0; ck < 5; ck++)
    for (int cj = 0; cj < 5; cj++)
      for (int ci = 0; ci < 5; ci++)
        for (int d = 0; d < 3; d++) {
          k_probe<<<nblocks(b0.nnodes), kBS, 0, s>>>(b0.X, b0.Y, b0.Z, ci, cj, ck, d, v);
          PTMF_TRY(matmult<double>(L, v, w));
          k_probe_scatter_dense<<<nblocks(b0.nnodes), kBS, 0, s>>>(b0.X, b0.Y, b0.Z, ci, cj, ck, d, w, mg->Adense);
          PTMF_HIP(hipGetLastError());
        }
  PTMF_HIP(hipStreamSynchronize(s));
  if (!mg->dinfo) PTMF_TRY(dalloc((void **)&mg->dinfo, sizeof(int)));
  rocblas_status st = rocsolver_dpotrf(mg->blas, rocblas_fill_lower, (int)n, mg->Adense, (int)n, mg->dinfo);
  PTMF_CHECK(st == rocblas_status_success, PTMF_ERR_DEVICE, "rocsolver_dpotrf: status %d", (int)st);
  int info = 0;
  PTMF_HIP(hipMemcpyAsync(&info, mg->dinfo, sizeof(int), hipMemcpyDeviceToHost, s));
  PTMF_HIP(hipStreamSynchronize(s));
  PTMF_CHECK(info == 0, PTMF_ERR_DIVERGED, "coarse Cholesky failed (info %d): operator not SPD", info);
  st = rocsolver_dpotri(mg->blas, rocblas_fill_lower, (int)n, mg->Adense, (int)n, mg->dinfo);
  PTMF_CHECK(st == rocblas_status_success, PTMF_ERR_DEVICE, "rocsolver_dpotri: status %d", (int)st);
  PTMF_HIP(hipMemcpyAsync(&info, mg->dinfo, sizeof(int), hipMemcpyDeviceToHost, s));
  PTMF_HIP(hipStreamSynchronize(s));
  PTMF_CHECK(info == 0, PTMF_ERR_DIVERGED, "coarse inverse failed (info %d)", info);
  // full symmetric inverse for the project's own product kernel (the V-cycle
  // graph then holds no library nodes); partial sums buffer of that product
  k_symmetrize<<<nblocks(n * n), kBS, 0, s>>>(n, mg->Adense);
  PTMF_HIP(hipGetLastError());
  if (!mg->dense_part)
    PTMF_TRY(dalloc((void **)&mg->dense_part, sizeof(double) * n * ((n + kGvCols - 1) / kGvCols)));
  PTMF_HIP(hipStreamSynchronize(s));
  return PTMF_OK;
}

// Level l's operator as its 125-point stencil (the layout of the Galerkin
// levels), probed through the level's own MatMult with the 375 colourings;
// Dirichlet dofs give unit rows and zero columns, as in the MatMult.
int probe_stencil(ptmf_mg_s *mg, int l, double *S) {
  MGLevel &L = mg->lev[l];
  hipStream_t s = mg->stream;
  const Box &b = L.op->box;
  const long N = b.nnodes;
  PTMF_HIP(hipMemsetAsync(S, 0, sizeof(double) * 1125 * N, s));
  double *v = L.p64[0], *w = L.r64;
  for (int ck = 0; ck < 5; ck++)
    for (int cj = 0; cj < 5; cj++)
      for (int ci = 0; ci < 5; ci++)
        for (int d = 0; d < 3; d++) {
          k_probe<<<nblocks(N), kBS, 0, s>>>(b.X, b.Y, b.Z, ci, cj, ck, d, v);
          PTMF_TRY(matmult<double>(L, v, w));
          k_probe_scatter<<<nblocks(N), kBS, 0, s>>>(b.X, b.Y, b.Z, ci, cj, ck, d, w, S);
          PTMF_HIP(hipGetLastError());
        }
  return PTMF_OK;
}

// Exact factor of level 0: the dense explicit inverse (at most kDenseMaxDofs)
// or the multifrontal Cholesky of its assembled stencil (any size).
int setup_coarse(ptmf_mg_s *mg) {
  MGLevel &L = mg->lev[0];
  mf_destroy(mg->mf);
  mg->mf = nullptr;
  const int cf = mg->opt.coarse_factor;
  if (cf == 1 || (cf == 0 && L.n <= kDenseMaxDofs)) return assemble_dense(mg);
  // the dense inverse of an earlier set-up must not be used any more
  (void)hipFree(mg->Adense);
  (void)hipFree(mg->dense_part);
  mg->Adense = mg->dense_part = nullptr;
  mg->ndense = 0;
  // all-Dirichlet nodes (identity rows, zero columns) stay out of the factor
  std::vector<uint8_t> bch(L.op->box.nnodes);
  PTMF_HIP(hipMemcpy(bch.data(), L.op->bcnode, L.op->box.nnodes, hipMemcpyDeviceToHost));
  if (L.galerkin) return mf_factor(L.op->box, L.S64, bch.data(), mg->stream, &mg->mf);
  double *S = nullptr;
  PTMF_TRY(dalloc((void **)&S, sizeof(double) * 1125 * L.op->box.nnodes));
  int st = probe_stencil(mg, 0, S);
  if (st == PTMF_OK) st = mf_factor(L.op->box, S, bch.data(), mg->stream, &mg->mf);
  (void)hipStreamSynchronize(mg->stream);
  (void)hipFree(S);
  return st;
}

void drop_one_graph(ptmf_mg_s::VGraph &v) {
  if (v.ex) (void)hipGraphExecDestroy(v.ex);
  if (v.g) (void)hipGraphDestroy(v.g);
  v = ptmf_mg_s::VGraph();
}

void drop_graph(ptmf_mg_s *mg) {
  for (auto &v : mg->vg) drop_one_graph(v);
  mg->vcalls = 0;
}

bool have_graph(const ptmf_mg_s *mg) {
  for (const auto &v : mg->vg)
    if (v.ex) return true;
  return false;
}

// One reduction of the low-synchronisation Gram-Schmidt step for column k:
// out[i] = <V_i, V_k> (i <= k), out[k + 1 + i] = <V_i, w> (i <= k),
// out[2k + 2] = <w, w>; one pass over V_0..V_k and w, one host synchronisation.
int lsgs_reduce(ptmf_mg_s *mg, long n, int k, const double *const *dV, const double *w, double *out) {
  PTMF_CHECK(k >= 0 && 2 * k + 3 <= kRedSlots, PTMF_ERR_ARG, "lsgs_reduce: column %d", k);
  const int nch = (k + 1 + 7) / 8;
  k_lsgs_partial<<<dim3(kRedBlocks, (unsigned)nch), kBS, 0, mg->stream>>>(n, k, dV, w, mg->red_part);
  k_mdot_final<<<1, kBS, 0, mg->stream>>>(2 * k + 3, kRedBlocks, mg->red_part, mg->red_out);
  PTMF_HIP(hipGetLastError());
  PTMF_HIP(hipMemcpyAsync(mg->red_host, mg->red_out, sizeof(double) * (2 * k + 3), hipMemcpyDeviceToHost,
                          mg->stream));
  PTMF_HIP(hipStreamSynchronize(mg->stream));
  memcpy(out, mg->red_host, sizeof(double) * (2 * k + 3));
  return PTMF_OK;
}

// Right-preconditioned flexible GMRES(restart) on vectors of length n with a
// low-synchronisation Gram-Schmidt process (one reduction and one update
// pass over the basis per iteration; the re-orthogonalisation of each basis
// vector is delayed to the next iteration's reduction, the DCGS2 scheme of
// Swirydowicz, Langou, Ananthan, Yang and Thomas, "Low synchronization
// Gram-Schmidt and generalized minimal residual algorithms", Numer. Linear
// Algebra Appl. 28 (2021) e2343).  At iteration k the one reduction gives
// <V_i, V_k>, <V_i, w> and <w, w> for w = A z_k; V_k (orthogonalised once,
// normalised by an estimate) is then orthogonalised again and normalised
// exactly, column k-1 of the Hessenberg matrix is corrected for that change
// and only then reduced by its Givens rotation, and w is orthogonalised
// against the corrected basis in the same update pass that writes V_k+1.
// So the residual estimate lags one iteration: at convergence one
// preconditioner application is not used.  apply(v, y): y = K v; prec(r, z):
// z ~ K^-1 r (may vary between calls).  The true residual is recomputed at
// every restart and at the end; stops when it (or the Arnoldi estimate)
// drops below rtol * ||b||.
template <typename Apply, typename Prec>
int fgmres_core(ptmf_mg_s *mg, KrylovSpace &ks, long n, Apply apply, Prec prec, const double *b, double *x,
                double rtol, int maxit, int restart, int *its_out, double *rnorm_out, double *history,
                int history_len) {
  hipStream_t s = mg->stream;
  if (ks.restart != restart || ks.n != n) {
    PTMF_HIP(hipStreamSynchronize(s));
    ks.release();
    ks.V.assign(restart + 1, nullptr);
    ks.Z.assign(restart, nullptr);
    for (auto &p : ks.V) PTMF_TRY(dalloc((void **)&p, sizeof(double) * n));
    for (auto &p : ks.Z) PTMF_TRY(dalloc((void **)&p, sizeof(double) * n));
    PTMF_TRY(dalloc((void **)&ks.w, sizeof(double) * n));
    PTMF_TRY(dalloc((void **)&ks.dV, sizeof(double *) * (restart + 1)));
    PTMF_TRY(dalloc((void **)&ks.dZ, sizeof(double *) * restart));
    PTMF_HIP(hipMemcpy((void *)ks.dV, ks.V.data(), sizeof(double *) * (restart + 1), hipMemcpyHostToDevice));
    PTMF_HIP(hipMemcpy((void *)ks.dZ, ks.Z.data(), sizeof(double *) * restart, hipMemcpyHostToDevice));
    ks.restart = restart;
    ks.n = n;
  }
  double *w = ks.w;
  std::vector<double> H((restart + 1) * restart), cs(restart), sn(restart), g(restart + 1), y(restart);
  std::vector<double> red(2 * restart + 3);
  double bnorm2;
  PTMF_TRY(dot(mg, n, b, b, &bnorm2));
  const double bnorm = sqrt(bnorm2);
  int its = 0, started = 0, hl = 0;
  double rnorm = 0.0;
  if (bnorm == 0.0) {
    PTMF_TRY(launch_fill<double>(x, 0.0, n, s));
    PTMF_HIP(hipStreamSynchronize(s));
    if (its_out) *its_out = 0;
    if (rnorm_out) *rnorm_out = 0.0;
    return PTMF_OK;
  }
  // Givens reduction of the (final) column j of H; returns |g_j+1|
  auto reduce_column = [&](int j) {
    double h[kMaxDots + 1];
    for (int i = 0; i <= j + 1; i++) h[i] = H[i * restart + j];
    for (int i = 0; i < j; i++) {
      const double t = cs[i] * h[i] + sn[i] * h[i + 1];
      h[i + 1] = -sn[i] * h[i] + cs[i] * h[i + 1];
      h[i] = t;
    }
    const double den = sqrt(h[j] * h[j] + h[j + 1] * h[j + 1]);
    cs[j] = den > 0 ? h[j] / den : 1.0;
    sn[j] = den > 0 ? h[j + 1] / den : 0.0;
    h[j] = den;
    h[j + 1] = 0.0;
    g[j + 1] = -sn[j] * g[j];
    g[j] = cs[j] * g[j];
    for (int i = 0; i <= j; i++) H[i * restart + j] = h[i];
    return fabs(g[j + 1]);
  };
  while (true) {
    PTMF_TRY(apply(x, w));
    k_residual<double><<<nblocks(n), kBS, 0, s>>>(n, b, w, ks.V[0]);
    PTMF_HIP(hipGetLastError());
    double rr;
    PTMF_TRY(dot(mg, n, ks.V[0], ks.V[0], &rr));
    rnorm = sqrt(rr);
    PTMF_CHECK(isfinite(rnorm), PTMF_ERR_DIVERGED, "fgmres: non-finite residual");
    if (history && hl < history_len) history[hl++] = rnorm;
    if (rnorm <= rtol * bnorm || started >= maxit) break;
    PTMF_TRY(scale<double>(s, n, 1.0 / rnorm, ks.V[0], ks.V[0]));
    std::fill(g.begin(), g.end(), 0.0);
    g[0] = rnorm;
    int ncol = 0;            // final columns of this cycle
    double eta_prev = 0.0;   // provisional h_k,k-1: the estimate V_k was normalised with
    for (int k = 0;; k++) {
      // extend by column k, or (end of cycle) only finalise column k-1
      const bool extend = k < restart && started < maxit;
      if (!extend && k == 0) break;
      const double *wv = w;
      if (extend) {
        PTMF_TRY(prec(ks.V[k], ks.Z[k]));
        PTMF_TRY(apply(ks.Z[k], w));
        started++;
      } else {
        wv = ks.V[k];
      }
      PTMF_TRY(lsgs_reduce(mg, n, k, ks.dV, wv, red.data()));
      const double *a = red.data(), *t = red.data() + (k + 1);
      const double ww = red[2 * k + 2];
      PTMF_CHECK(isfinite(ww) && isfinite(a[k]), PTMF_ERR_DIVERGED, "fgmres: non-finite Arnoldi vector");
      double rho = 1.0;
      if (k > 0) {
        // V_k = rho V_k' + sum_i<k a_i V_i: column k-1 in the corrected basis
        double aa = 0.0;
        for (int i = 0; i < k; i++) aa += a[i] * a[i];
        const double r2 = a[k] - aa;
        rho = r2 > 1e-30 * a[k] ? sqrt(r2) : 0.0;  // 0: A z_k-1 lies in the basis (breakdown)
        for (int i = 0; i < k; i++) H[i * restart + k - 1] += eta_prev * a[i];
        H[k * restart + k - 1] = eta_prev * rho;
        const double est = reduce_column(k - 1);
        its++;
        ncol = k;
        if (history && hl < history_len) history[hl++] = est;
        if (est <= rtol * bnorm || rho == 0.0) break;
      }
      if (!extend) break;
      // column k against the corrected basis: t_k' = <V_k', w>
      double tk = t[k];
      for (int i = 0; i < k; i++) tk -= a[i] * t[i];
      tk = rho > 0.0 ? tk / rho : 0.0;
      double e2 = ww - tk * tk;
      for (int i = 0; i < k; i++) {
        H[i * restart + k] = t[i];
        e2 -= t[i] * t[i];
      }
      H[k * restart + k] = tk;
      // the norm of the orthogonalised w by Pythagoras; only a scale: the
      // next reduction measures V_k+1 exactly and corrects column k
      double eta = e2 > 1e-14 * ww ? sqrt(e2) : (ww > 0.0 ? 1e-7 * sqrt(ww) : 1.0);
      double *c = mg->coef_host;
      for (int i = 0; i < k; i++) {
        c[i] = a[i];
        c[k + i] = t[i];
      }
      c[2 * k] = tk;
      c[2 * k + 1] = k > 0 ? 1.0 / rho : 1.0;
      c[2 * k + 2] = 1.0 / eta;
      PTMF_HIP(hipMemcpyAsync(mg->coef_dev, c, sizeof(double) * (2 * k + 3), hipMemcpyHostToDevice, s));
      k_lsgs_update<<<nblocks(n), kBS, 0, s>>>(n, k, (double *const *)ks.dV, mg->coef_dev, w);
      PTMF_HIP(hipGetLastError());
      // the coefficient buffer is reused by the next iteration
      PTMF_HIP(hipStreamSynchronize(s));
      eta_prev = eta;
    }
    // x += Z y over the final columns
    const int j = ncol;
    if (j > 0) {
      for (int i = j - 1; i >= 0; i--) {
        double tt = g[i];
        for (int k = i + 1; k < j; k++) tt -= H[i * restart + k] * y[k];
        y[i] = tt / H[i * restart + i];
      }
      memcpy(mg->coef_host, y.data(), sizeof(double) * j);
      PTMF_HIP(hipMemcpyAsync(mg->coef_dev, mg->coef_host, sizeof(double) * j, hipMemcpyHostToDevice, s));
      k_mupdate<<<nblocks(n), kBS, 0, s>>>(n, j, ks.dZ, mg->coef_dev, x);
      PTMF_HIP(hipGetLastError());
      PTMF_HIP(hipStreamSynchronize(s));
    }
  }
  PTMF_HIP(hipStreamSynchronize(s));
  if (its_out) *its_out = its;
  if (rnorm_out) *rnorm_out = rnorm;
  return PTMF_OK;
}

}  // namespace
}  // namespace ptmf
extern "C" int ptmf_mg_fcg(ptmf_mg mg, const double *b, double *x, double rtol, int maxit, int *its_out,
                           double *rnorm_out, double *history, int history_len);
namespace ptmf {
namespace {

// Saddle point K = [A11 A12; A21 0] on [u | p] (3N + 4E values):
// yu = A11 xu + A12 xp with Dirichlet rows y = x, yp = A21 xu
// (MatMult of the reference's MatNest Stokes operator, stokes_operators.c).
// y = K x in one pass over the elements (the fused saddle-point kernel)
int stokes_apply(ptmf_mg_s *mg, const double *x, double *y) {
  return launch_stokes_mult(mg->lev[mg->L - 1].op, x, y);
}

// One V-cycle from b into z for callers whose vectors change from call to
// call (FGMRES hands the preconditioner a new basis vector every iteration):
// the cycle runs between two buffers of the hierarchy's own, b copied in
// unless it is already the fixed input and the result copied out, so one
// captured graph serves every iteration (a graph is keyed by its pointers).
int vcycle_fixed(ptmf_mg_s *mg, const double *b, double *z) {
  const long n = mg->lev[mg->L - 1].n;
  if (!mg->kvb) PTMF_TRY(dalloc((void **)&mg->kvb, sizeof(double) * n));
  if (!mg->kvx) PTMF_TRY(dalloc((void **)&mg->kvx, sizeof(double) * n));
  const double *in = b;
  if (b != mg->stmp) {
    PTMF_HIP(hipMemcpyAsync(mg->kvb, b, sizeof(double) * n, hipMemcpyDeviceToDevice, mg->stream));
    in = mg->kvb;
  }
  PTMF_TRY(ptmf_mg_vcycle(mg, in, mg->kvx));
  PTMF_HIP(hipMemcpyAsync(z, mg->kvx, sizeof(double) * n, hipMemcpyDeviceToDevice, mg->stream));
  return PTMF_OK;
}

// Upper block-triangular preconditioner [A11 A12; 0 S]^-1 with S the scaled
// viscous pressure mass (block diagonal) and A11^-1 one V-cycle:
// zp = S^-1 rp, zu = V-cycle(ru - A12 zp)  (the fieldsplit Schur "upper"
// factorisation of the reference's solver configuration).
int stokes_prec(ptmf_mg_s *mg, const double *r, double *z) {
  MGLevel &F = mg->lev[mg->L - 1];
  const long nu = F.n;
  const long np = 4 * F.op->box.nel;
  PTMF_TRY(launch_schur_apply(F.op, mg->sMinv, r + nu, z + nu));
  // stmp = r_u - A12 z_p, A12 z_p from the fused saddle-point kernel on
  // [0 | z_p] (one element pass; its Dirichlet rows come out as 0).  With an
  // FP32 velocity V-cycle the product is formed in FP32 like the rest of the
  // velocity preconditioner (its result only feeds that V-cycle); z_p, the
  // Schur update, stays FP64.
  if (mg->opt.smoother_precision == 32 && mg->sx2f) {
    PTMF_TRY(launch_fill<float>(mg->sx2f, 0.0f, nu, mg->stream));
    k_cast<double, float><<<nblocks(np), kBS, 0, mg->stream>>>(np, z + nu, mg->sx2f + nu);
    PTMF_HIP(hipGetLastError());
    PTMF_TRY(launch_stokes_mult_f32(F.op, mg->sx2f, mg->sy2f));
    k_residual_mixed<<<nblocks(nu), kBS, 0, mg->stream>>>(nu, r, mg->sy2f, mg->stmp);
  } else {
    PTMF_TRY(launch_fill<double>(mg->sx2, 0.0, nu, mg->stream));
    PTMF_HIP(hipMemcpyAsync(mg->sx2 + nu, z + nu, sizeof(double) * np, hipMemcpyDeviceToDevice, mg->stream));
    PTMF_TRY(launch_stokes_mult(F.op, mg->sx2, mg->sy2));
    k_residual<double><<<nblocks(nu), kBS, 0, mg->stream>>>(nu, r, mg->sy2, mg->stmp);
  }
  PTMF_HIP(hipGetLastError());
  if (mg->s_inner_its <= 1) return vcycle_fixed(mg, mg->stmp, z);
  PTMF_HIP(hipMemsetAsync(z, 0, sizeof(double) * nu, mg->stream));
  int its = 0;
  double rn = 0.0;
  return ptmf_mg_fcg(mg, mg->stmp, z, mg->s_inner_rtol, mg->s_inner_its, &its, &rn, nullptr, 0);
}

void set_op_stream(ptmf_op op, hipStream_t s) {
  if (op->stream && op->stream != s) (void)hipStreamDestroy(op->stream);
  op->stream = s;
}

}  // namespace
}  // namespace ptmf

using ptmf::MGLevel;

extern "C" {

int ptmf_mg_default_options(ptmf_mg_options *o) {
  PTMF_CHECK(o, PTMF_ERR_ARG, "mg_default_options: null");
  o->levels_mg = 0;
  o->smoother_precision = 32;
  o->pre_its = 3;
  o->post_its = 3;
  o->esteig[0] = 0.0;
  o->esteig[1] = 0.2;
  o->esteig[2] = 0.0;
  o->esteig[3] = 1.1;
  o->esteig_its = 20;
  o->coarse_rtol = 1e-3;
  o->coarse_maxit = 200;
  o->galerkin_levels = 0;
  o->bc_transfer = 1;
  o->transfer = 0;
  o->smoother_diag = 0;
  o->use_graph = 1;
  o->coarse_factor = 0;
  return PTMF_OK;
}

int ptmf_mg_create(int mx, int my, int mz, int nlevels, ptmf_mg *out) {
  PTMF_CHECK(out, PTMF_ERR_ARG, "mg_create: null");
  *out = nullptr;
  PTMF_CHECK(nlevels >= 1 && nlevels <= 16, PTMF_ERR_ARG, "mg_create: nlevels %d (1..16)", nlevels);
  for (int l = 1; l < nlevels; l++) {
    const int f = 1 << l;
    PTMF_CHECK(mx % f == 0 && my % f == 0 && mz % f == 0, PTMF_ERR_ARG,
               "mg_create: %d x %d x %d cannot be coarsened %d times by 2", mx, my, mz, nlevels - 1);
  }
  ptmf_mg_s *mg = new ptmf_mg_s();
  ptmf_mg_default_options(&mg->opt);
  mg->L = nlevels;
  int st = PTMF_OK;
  hipError_t he = hipStreamCreate(&mg->stream);
  if (he != hipSuccess) st = ptmf_set_error(PTMF_ERR_DEVICE, "mg_create: %s", hipGetErrorString(he));
  for (int l = 0; l < nlevels && st == PTMF_OK; l++) {
    MGLevel &L = mg->lev[l];
    const int f = 1 << (nlevels - 1 - l);
    L.mx = mx / f; L.my = my / f; L.mz = mz / f;
    st = ptmf_op_create(L.mx, L.my, L.mz, &L.op);
    if (st != PTMF_OK) break;
    ptmf::set_op_stream(L.op, mg->stream);
    L.n = 3 * L.op->box.nnodes;
    const size_t b8 = sizeof(double) * L.n, b4 = sizeof(float) * L.n;
    void **a8[] = {(void **)&L.diag, (void **)&L.dinv64, (void **)&L.b64, (void **)&L.x64,
                   (void **)&L.r64, (void **)&L.p64[0], (void **)&L.p64[1], (void **)&L.p64[2]};
    for (void **p : a8)
      if (st == PTMF_OK) st = ptmf::dalloc(p, b8);
    void **a4[] = {(void **)&L.dinv32, (void **)&L.b32, (void **)&L.x32, (void **)&L.r32,
                   (void **)&L.p32[0], (void **)&L.p32[1], (void **)&L.p32[2]};
    for (void **p : a4)
      if (st == PTMF_OK) st = ptmf::dalloc(p, b4);
  }
  if (st == PTMF_OK) st = ptmf::dalloc((void **)&mg->red_part, sizeof(double) * ptmf::kRedBlocks * ptmf::kRedSlots);
  if (st == PTMF_OK) st = ptmf::dalloc((void **)&mg->red_out, sizeof(double) * ptmf::kRedSlots);
  if (st == PTMF_OK) st = ptmf::dalloc((void **)&mg->coef_dev, sizeof(double) * ptmf::kRedSlots);
  if (st == PTMF_OK) {
    he = hipHostMalloc((void **)&mg->red_host, sizeof(double) * ptmf::kRedSlots, hipHostMallocDefault);
    if (he == hipSuccess)
      he = hipHostMalloc((void **)&mg->coef_host, sizeof(double) * ptmf::kRedSlots, hipHostMallocDefault);
    if (he != hipSuccess) st = ptmf_set_error(PTMF_ERR_NOMEM, "mg_create: pinned buffer: %s", hipGetErrorString(he));
  }
  if (st == PTMF_OK) {
    rocblas_status bs = rocblas_create_handle(&mg->blas);
    if (bs != rocblas_status_success) st = ptmf_set_error(PTMF_ERR_DEVICE, "rocblas_create_handle: %d", (int)bs);
    else rocblas_set_stream(mg->blas, mg->stream);
  }
  if (st != PTMF_OK) {
    ptmf_mg tmp = mg;
    ptmf_mg_destroy(&tmp);
    return st;
  }
  *out = mg;
  return PTMF_OK;
}

int ptmf_mg_destroy(ptmf_mg *pmg) {
  if (!pmg || !*pmg) return PTMF_OK;
  ptmf_mg_s *mg = *pmg;
  if (mg->stream) (void)hipStreamSynchronize(mg->stream);
  for (int l = 0; l < mg->L; l++) {
    MGLevel &L = mg->lev[l];
    if (L.op) {
      L.op->stream = nullptr;  // shared stream, destroyed below
      ptmf_op_destroy(&L.op);
    }
    void *ps[] = {L.diag, L.dinv64, L.b64, L.x64, L.r64, L.p64[0], L.p64[1], L.p64[2],
                  L.dinv32, L.b32, L.x32, L.r32, L.p32[0], L.p32[1], L.p32[2], L.S64, L.S32};
    for (void *p : ps) (void)hipFree(p);
  }
  mg->ku.release();
  mg->ks.release();
  (void)hipFree(mg->sMinv);
  (void)hipFree(mg->stmp);
  (void)hipFree(mg->sx2);
  (void)hipFree(mg->sy2);
  (void)hipFree(mg->sx2f);
  (void)hipFree(mg->sy2f);
  (void)hipFree(mg->kvb);
  (void)hipFree(mg->kvx);
  ptmf::drop_graph(mg);
  for (double *p : mg->pcg) (void)hipFree(p);
  for (double *p : mg->fcg) (void)hipFree(p);
  (void)hipFree(mg->red_part);
  (void)hipFree(mg->red_out);
  (void)hipFree(mg->coef_dev);
  if (mg->coef_host) (void)hipHostFree(mg->coef_host);
  ptmf::mf_destroy(mg->mf);
  (void)hipFree(mg->Adense);
  (void)hipFree(mg->dense_part);
  (void)hipFree(mg->dinfo);
  if (mg->red_host) (void)hipHostFree(mg->red_host);
  if (mg->blas) rocblas_destroy_handle(mg->blas);
  if (mg->stream) (void)hipStreamDestroy(mg->stream);
  delete mg;
  *pmg = nullptr;
  return PTMF_OK;
}

int ptmf_mg_levels(ptmf_mg mg, int *nlevels) {
  PTMF_CHECK(mg && nlevels, PTMF_ERR_ARG, "mg_levels: null");
  *nlevels = mg->L;
  return PTMF_OK;
}

int ptmf_mg_level_op(ptmf_mg mg, int level, ptmf_op *op) {
  PTMF_CHECK(mg && op && level >= 0 && level < mg->L, PTMF_ERR_ARG, "mg_level_op: bad level %d", level);
  *op = mg->lev[level].op;
  return PTMF_OK;
}

int ptmf_mg_set_coordinates(ptmf_mg mg, const double *coords) {
  PTMF_CHECK(mg && coords, PTMF_ERR_ARG, "mg_set_coordinates: null");
  // finest level takes the input; coarser levels by injection
  // (DMDARestrictCoordinatesHierarchy, dmda_project_coords.c:39-71)
  std::vector<double> cur(coords, coords + 3 * mg->lev[mg->L - 1].op->box.nnodes);
  for (int l = mg->L - 1; l >= 0; l--) {
    MGLevel &L = mg->lev[l];
    PTMF_TRY(ptmf_op_set_coordinates(L.op, cur.data()));
    if (l == 0) break;
    const ptmf::Box &bf = L.op->box, &bc = mg->lev[l - 1].op->box;
    std::vector<double> nxt(3 * bc.nnodes);
    for (long k = 0; k < bc.Z; k++)
      for (long j = 0; j < bc.Y; j++)
        for (long i = 0; i < bc.X; i++) {
          const long g = i + bc.X * (j + bc.Y * k), f = 2 * i + bf.X * (2 * j + bf.Y * 2 * k);
          for (int d = 0; d < 3; d++) nxt[3 * g + d] = cur[3 * f + d];
        }
    cur.swap(nxt);
  }
  mg->is_setup = false;
  return PTMF_OK;
}

int ptmf_mg_set_viscosity(ptmf_mg mg, int level, const double *eta) {
  PTMF_CHECK(mg && eta && level >= 0 && level < mg->L, PTMF_ERR_ARG, "mg_set_viscosity: bad level %d", level);
  PTMF_TRY(ptmf_op_set_viscosity(mg->lev[level].op, eta));
  mg->lev[level].have_eta = true;
  mg->is_setup = false;
  return PTMF_OK;
}

int ptmf_mg_coarsen_viscosity(ptmf_mg mg, int mode) {
  PTMF_CHECK(mg, PTMF_ERR_ARG, "mg_coarsen_viscosity: null");
  PTMF_CHECK(mode >= 0 && mode <= 6, PTMF_ERR_UNSUPPORTED,
             "mg_coarsen_viscosity: mode %d (0 arithmetic, 1 geometric, 2 harmonic element means; "
             "3 arithmetic, 4 geometric quadrature-point-local means; 5 Q1 projection with scaled "
             "restriction, 6 with injection)", mode);
  PTMF_CHECK(mg->lev[mg->L - 1].have_eta, PTMF_ERR_NOT_SETUP, "mg_coarsen_viscosity: finest viscosity not set");
  // device scratch: coarse element-major values and, for the projection, two
  // nodal fields (fine and next coarser)
  const ptmf::MGLevel &Fin = mg->lev[mg->L - 1];
  const long nnf = (long)(2 * Fin.mx + 1) * (2 * Fin.my + 1) * (2 * Fin.mz + 1);
  double *ec = nullptr, *nod[2] = {nullptr, nullptr};
  int st = ptmf::dalloc((void **)&ec, sizeof(double) * 27 * (Fin.op->box.nel / 8 + 1));
  if (st == PTMF_OK && mode >= 5) st = ptmf::dalloc((void **)&nod[0], sizeof(double) * nnf);
  if (st == PTMF_OK && mode >= 5) st = ptmf::dalloc((void **)&nod[1], sizeof(double) * nnf);
  hipStream_t s = mg->stream;
  if (st == PTMF_OK && mode >= 5) {
    ptmf::k_q1_nodal<<<ptmf::nblocks(nnf), ptmf::kBS, 0, s>>>(Fin.mx, Fin.my, Fin.mz, Fin.op->d64.eta, nod[0]);
    if (hipGetLastError() != hipSuccess) st = ptmf_set_error(PTMF_ERR_DEVICE, "mg_coarsen_viscosity: launch");
  }
  int cur = 0;
  for (int l = mg->L - 1; l > 0 && st == PTMF_OK; l--) {
    ptmf::MGLevel &F = mg->lev[l], &Cl = mg->lev[l - 1];
    const long nec = Cl.op->box.nel;
    if (mode <= 2) {
      ptmf::k_coarsen_eta<<<ptmf::nblocks(nec), ptmf::kBS, 0, s>>>(Cl.mx, Cl.my, Cl.mz, mode, F.op->d64.eta, ec);
    } else if (mode <= 4) {
      ptmf::k_coarsen_eta_qp<<<ptmf::nblocks(27 * nec), ptmf::kBS, 0, s>>>(Cl.mx, Cl.my, Cl.mz, mode,
                                                                         F.op->d64.eta, ec);
    } else {
      const long nnc = (long)(2 * Cl.mx + 1) * (2 * Cl.my + 1) * (2 * Cl.mz + 1);
      ptmf::k_q1_restrict<<<ptmf::nblocks(nnc), ptmf::kBS, 0, s>>>(Cl.mx, Cl.my, Cl.mz, mode == 6, nod[cur],
                                                                  nod[1 - cur]);
      cur = 1 - cur;
      ptmf::k_q1_to_qp<<<ptmf::nblocks(27 * nec), ptmf::kBS, 0, s>>>(Cl.mx, Cl.my, Cl.mz, nod[cur], ec);
    }
    if (hipGetLastError() != hipSuccess) {
      st = ptmf_set_error(PTMF_ERR_DEVICE, "mg_coarsen_viscosity: launch");
      break;
    }
    // the coarse operator's qp-major FP64 and FP32 viscosity, written on the device
    st = ptmf::launch_eta_qpmajor(nec, ec, Cl.op->d64.eta, Cl.op->d32.eta, s);
    if (st == PTMF_OK && hipStreamSynchronize(s) != hipSuccess)
      st = ptmf_set_error(PTMF_ERR_DEVICE, "mg_coarsen_viscosity: synchronize");
    if (st == PTMF_OK) {
      Cl.op->has_eta = 1;
      Cl.op->state++;
      Cl.have_eta = true;
    }
  }
  (void)hipStreamSynchronize(s);
  (void)hipFree(ec);
  (void)hipFree(nod[0]);
  (void)hipFree(nod[1]);
  if (st != PTMF_OK) return st;
  mg->is_setup = false;
  return PTMF_OK;
}

int ptmf_mg_set_dirichlet(ptmf_mg mg, int level, const unsigned char *mask) {
  PTMF_CHECK(mg && mask && level >= 0 && level < mg->L, PTMF_ERR_ARG, "mg_set_dirichlet: bad level %d", level);
  PTMF_TRY(ptmf_op_set_dirichlet(mg->lev[level].op, mask));
  mg->lev[level].have_bc = true;
  mg->is_setup = false;
  return PTMF_OK;
}

int ptmf_mg_setup(ptmf_mg mg, const ptmf_mg_options *opt) {
  PTMF_CHECK(mg, PTMF_ERR_ARG, "mg_setup: null");
  ptmf::drop_graph(mg);
  if (opt) mg->opt = *opt;
  ptmf_mg_options &o = mg->opt;
  if (o.levels_mg <= 0 || o.levels_mg > mg->L) o.levels_mg = mg->L;
  PTMF_CHECK(o.smoother_precision == 32 || o.smoother_precision == 64, PTMF_ERR_UNSUPPORTED,
             "mg_setup: smoother precision %d (32 or 64)", o.smoother_precision);
  PTMF_CHECK(o.pre_its >= 0 && o.post_its >= 0 && o.esteig_its > 0, PTMF_ERR_ARG, "mg_setup: bad iteration counts");
  PTMF_CHECK(o.transfer == 0 || o.transfer == 1, PTMF_ERR_UNSUPPORTED, "mg_setup: transfer %d (0 or 1)",
             o.transfer);
  PTMF_CHECK(o.smoother_diag == 0 || o.smoother_diag == 1, PTMF_ERR_UNSUPPORTED,
             "mg_setup: smoother_diag %d (0 Jacobi, 1 l1-Jacobi)", o.smoother_diag);
  PTMF_CHECK(o.galerkin_levels >= 0 && o.galerkin_levels < mg->L, PTMF_ERR_ARG,
             "mg_setup: galerkin_levels %d (0..nlevels-1)", o.galerkin_levels);
  hipStream_t s = mg->stream;
  // Galerkin coarse operators, built top-down from the finer level's operator
  for (int l = mg->L - 1; l >= 0; l--) mg->lev[l].galerkin = false;
  for (int l = mg->L - 2; l >= 0; l--) {
    MGLevel &L = mg->lev[l];
    PTMF_CHECK(L.op->has_coords && L.have_bc, PTMF_ERR_NOT_SETUP,
               "mg_setup: level %d lacks coordinates or Dirichlet list", l);
    if (l < o.galerkin_levels) PTMF_TRY(ptmf::assemble_galerkin(mg, l));
  }
  for (int l = 0; l < mg->L; l++) {
    MGLevel &L = mg->lev[l];
    PTMF_CHECK(L.op->has_coords && (L.have_eta || L.galerkin) && L.have_bc, PTMF_ERR_NOT_SETUP,
               "mg_setup: level %d lacks coordinates, viscosity or Dirichlet list", l);
    if (L.galerkin) {
      if (o.smoother_diag == 1)
        ptmf::k_stencil_l1<<<ptmf::nblocks(L.op->box.nnodes), ptmf::kBS, 0, s>>>(L.op->box.nnodes, L.S64, L.diag);
      else
        ptmf::k_stencil_diag<<<ptmf::nblocks(L.op->box.nnodes), ptmf::kBS, 0, s>>>(L.op->box.nnodes, L.S64, L.diag);
      PTMF_HIP(hipGetLastError());
    } else if (o.smoother_diag == 1) {
      PTMF_TRY(ptmf_a11_l1_diagonal(L.op, L.diag));
    } else {
      PTMF_TRY(ptmf_a11_diagonal(L.op, L.diag));
    }
    ptmf::k_recip<double><<<ptmf::nblocks(L.n), ptmf::kBS, 0, s>>>(L.n, L.diag, L.dinv64);
    ptmf::k_recip<float><<<ptmf::nblocks(L.n), ptmf::kBS, 0, s>>>(L.n, L.diag, L.dinv32);
    PTMF_HIP(hipGetLastError());
    if (l > 0) {
      double est = 0.0;
      PTMF_TRY(ptmf::estimate_emax(mg, l, o.esteig_its, &est));
      // KSPChebyshevEstEigSet(ksp, a, b, c, d): emin = a*min + b*max, emax = c*min + d*max;
      // with the default transform only the max estimate enters
      L.emin = o.esteig[0] * est + o.esteig[1] * est;
      L.emax = o.esteig[2] * est + o.esteig[3] * est;
      PTMF_TRY(ptmf::place_dirichlet(mg, L));
    }
  }
  PTMF_CHECK(o.coarse_factor >= 0 && o.coarse_factor <= 2, PTMF_ERR_UNSUPPORTED,
             "mg_setup: coarse_factor %d (0 auto, 1 dense, 2 sparse)", o.coarse_factor);
  PTMF_TRY(ptmf::setup_coarse(mg));
  // coarse PCG workspace (level levels_mg coarse index)
  const int lc = mg->L - o.levels_mg;
  if (lc > 0) {
    const long n = mg->lev[lc].n;
    for (int i = 0; i < 4; i++) {
      if (mg->pcg[i]) (void)hipFree(mg->pcg[i]);
      PTMF_TRY(ptmf::dalloc((void **)&mg->pcg[i], sizeof(double) * n));
    }
  }
  PTMF_HIP(hipStreamSynchronize(s));
  mg->is_setup = true;
  return PTMF_OK;
}

int ptmf_mg_coarse_info(ptmf_mg mg, int *kind, long *bytes, int *depths, double *setup_ms) {
  PTMF_CHECK(mg, PTMF_ERR_ARG, "mg_coarse_info: null");
  PTMF_CHECK(mg->is_setup, PTMF_ERR_NOT_SETUP, "mg_coarse_info: not set up");
  const bool sparse = mg->mf != nullptr;
  if (kind) *kind = sparse ? 2 : 1;
  if (bytes) *bytes = sparse ? 8 * ptmf::mf_factor_doubles(mg->mf) : 8L * mg->ndense * mg->ndense;
  if (depths) *depths = sparse ? ptmf::mf_depths(mg->mf) : 0;
  if (setup_ms) *setup_ms = sparse ? ptmf::mf_setup_ms(mg->mf) : 0.0;
  return PTMF_OK;
}

int ptmf_mg_level_eigs(ptmf_mg mg, int level, double *emin, double *emax) {
  PTMF_CHECK(mg && level >= 0 && level < mg->L, PTMF_ERR_ARG, "mg_level_eigs: bad level");
  if (emin) *emin = mg->lev[level].emin;
  if (emax) *emax = mg->lev[level].emax;
  return PTMF_OK;
}

int ptmf_mg_set_level_eigs(ptmf_mg mg, int level, double emin, double emax) {
  PTMF_CHECK(mg && level >= 0 && level < mg->L && emax > emin && emin > 0, PTMF_ERR_ARG,
             "mg_set_level_eigs: bad arguments");
  mg->lev[level].emin = emin;
  mg->lev[level].emax = emax;
  ptmf::drop_graph(mg);
  if (mg->is_setup) PTMF_TRY(ptmf::place_dirichlet(mg, mg->lev[level]));
  return PTMF_OK;
}

// x = V-cycle(b) on the finest level (FP64 in and out)
namespace ptmf {
namespace {
int vcycle_body(ptmf_mg_s *mg, const double *b, double *x) {
  const int lf = mg->L - 1, lc = mg->L - mg->opt.levels_mg;
  MGLevel &F = mg->lev[lf];
  if (mg->opt.smoother_precision == 64 || lf == lc) return vcycle<double>(mg, lf, lc, b, x);
  return vcycle<float>(mg, lf, lc, F.b32, F.x32, b, x);
}

void graph_key(const ptmf_op_s *op, ptmf_mg_s::GraphKey *out) {
  ptmf_mg_s::GraphKey &k = *out;
  k.state = op->state;
  k.TX = op->tiling.TX;
  k.TY = op->tiling.TY;
  k.ZC = op->tiling.ZC;
  k.variant = op->variant;
  k.wg = op->wg_tiles;
  k.faces = op->faces;
  k.bcidx = op->bcidx;
  k.et64 = op->et64;
  k.et32 = op->et32;
  k.nbc = op->nbc;
}

// Capture one V-cycle from b into x as a HIP graph (into the least recently
// used cache slot).  Everything the cycle builds lazily (tile-ordered data)
// exists after the first, eager call.
int capture_vcycle(ptmf_mg_s *mg, const double *b, double *x, ptmf_mg_s::VGraph **out) {
  ptmf_mg_s::VGraph *slot = &mg->vg[0];
  for (auto &v : mg->vg)
    if (!v.ex || v.used < slot->used) {
      slot = &v;
      if (!v.ex) break;
    }
  drop_one_graph(*slot);
  PTMF_HIP(hipStreamSynchronize(mg->stream));
  PTMF_HIP(hipStreamBeginCapture(mg->stream, hipStreamCaptureModeThreadLocal));
  const int st = vcycle_body(mg, b, x);
  hipGraph_t g = nullptr;
  const hipError_t ce = hipStreamEndCapture(mg->stream, &g);
  if (st != PTMF_OK) {
    if (g) (void)hipGraphDestroy(g);
    return st;
  }
  PTMF_CHECK(ce == hipSuccess && g, PTMF_ERR_DEVICE, "mg_vcycle: graph capture failed: %s", hipGetErrorString(ce));
  hipGraphExec_t ex = nullptr;
  const hipError_t ie = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
  if (ie != hipSuccess) {
    (void)hipGraphDestroy(g);
    return ptmf_set_error(PTMF_ERR_DEVICE, "mg_vcycle: graph instantiation failed: %s", hipGetErrorString(ie));
  }
  slot->b = b;
  slot->x = x;
  slot->g = g;
  slot->ex = ex;
  for (int l = 0; l < mg->L; l++) graph_key(mg->lev[l].op, &mg->gkey[l]);
  *out = slot;
  return PTMF_OK;
}

bool graph_stale(const ptmf_mg_s *mg) {
  for (int l = 0; l < mg->L; l++) {
    ptmf_mg_s::GraphKey k;
    graph_key(mg->lev[l].op, &k);
    const ptmf_mg_s::GraphKey &c = mg->gkey[l];
    if (k.state != c.state || k.TX != c.TX || k.TY != c.TY || k.ZC != c.ZC || k.variant != c.variant || k.wg != c.wg ||
        k.faces != c.faces || k.bcidx != c.bcidx || k.et64 != c.et64 || k.et32 != c.et32 || k.nbc != c.nbc)
      return true;
  }
  return false;
}
}  // namespace
}  // namespace ptmf

// x = V-cycle(b) on the finest level (FP64 in and out).  With use_graph (and
// the whole hierarchy in the V-cycle) the second and later calls replay a
// captured graph.
int ptmf_mg_vcycle(ptmf_mg mg, const double *b, double *x) {
  PTMF_CHECK(mg && b && x, PTMF_ERR_ARG, "mg_vcycle: null");
  PTMF_CHECK(mg->is_setup, PTMF_ERR_NOT_SETUP, "mg_vcycle: not set up");
  const bool graph = mg->opt.use_graph && mg->opt.levels_mg == mg->L;
  // a level operator changed since the capture: replay nothing stale, run
  // the next cycle eagerly (it rebuilds the operator's derived data) and
  // capture again after it
  if (graph && ptmf::have_graph(mg) && ptmf::graph_stale(mg)) ptmf::drop_graph(mg);
  if (!graph || mg->vcalls++ == 0) return ptmf::vcycle_body(mg, b, x);
  ptmf_mg_s::VGraph *v = nullptr;
  for (auto &c : mg->vg)
    if (c.ex && c.b == b && c.x == x) v = &c;
  if (!v) PTMF_TRY(ptmf::capture_vcycle(mg, b, x, &v));
  v->used = ++mg->vtick;
  PTMF_HIP(hipGraphLaunch(v->ex, mg->stream));
  return PTMF_OK;
}

int ptmf_mg_level_apply(ptmf_mg mg, int level, int precision, const void *x, void *y) {
  PTMF_CHECK(mg && x && y && level >= 0 && level < mg->L, PTMF_ERR_ARG, "mg_level_apply: bad level %d", level);
  PTMF_CHECK(mg->is_setup, PTMF_ERR_NOT_SETUP, "mg_level_apply: not set up");
  if (precision == 64)
    PTMF_TRY(ptmf::matmult<double>(mg->lev[level], (const double *)x, (double *)y));
  else if (precision == 32)
    PTMF_TRY(ptmf::matmult<float>(mg->lev[level], (const float *)x, (float *)y));
  else
    return ptmf_set_error(PTMF_ERR_UNSUPPORTED, "mg_level_apply: precision %d", precision);
  PTMF_HIP(hipStreamSynchronize(mg->stream));
  return PTMF_OK;
}

// Level operator as a CSR matrix on the host (MatAssemble_StokesA_AUU,
// stokes_assembly.c:115-297, with BCListInsertScaling, dmda_bcs.c:1035-1060):
// rows and columns in the node-interlaced dof order 3n + c; a Dirichlet row
// holds its unit diagonal only, a free row the couplings to the free dofs of
// every node that shares an element with its node, columns ascending.  The
// values come from the level's 125-point stencil (its Galerkin stencil, or the
// rediscretised operator probed through its own MatMult with the 375
// colourings).  With colidx == vals == NULL only *nnz (and rowptr, when not
// NULL) is computed.
int ptmf_mg_level_assemble(ptmf_mg mg, int level, long *nnz_out, long *rowptr, long *colidx, double *vals) {
  PTMF_CHECK(mg && nnz_out && level >= 0 && level < mg->L, PTMF_ERR_ARG, "mg_level_assemble: bad arguments");
  PTMF_CHECK((colidx == nullptr) == (vals == nullptr), PTMF_ERR_ARG,
             "mg_level_assemble: colidx and vals are given together");
  PTMF_CHECK(!colidx || rowptr, PTMF_ERR_ARG, "mg_level_assemble: the fill call needs rowptr");
  PTMF_CHECK(mg->is_setup, PTMF_ERR_NOT_SETUP, "mg_level_assemble: not set up");
  ptmf::MGLevel &L = mg->lev[level];
  const ptmf::Box &b = L.op->box;
  const long X = b.X, Y = b.Y, Z = b.Z, N = b.nnodes, XY = X * Y;
  std::vector<uint8_t> bc(N);
  PTMF_HIP(hipMemcpy(bc.data(), L.op->bcnode, N, hipMemcpyDeviceToHost));
  // nodes a and a + o of one lattice direction (extent E) share an element
  // when |o| <= 1, or |o| == 2 with both indices even (element boundaries)
  auto couples = [](long a, int o, long E) {
    const long c = a + o;
    if (c < 0 || c >= E) return false;
    return o >= -1 && o <= 1 ? true : (a % 2) == 0;
  };
  // pass 1: row lengths
  long nnz = 0;
  if (rowptr) rowptr[0] = 0;
  for (long m = 0; m < N; m++) {
    const long i = m % X, j = (m / X) % Y, k = m / XY;
    long nfree = 0;  // free dofs of the nodes coupled to m
    for (int ok = -2; ok <= 2; ok++) {
      if (!couples(k, ok, Z)) continue;
      for (int oj = -2; oj <= 2; oj++) {
        if (!couples(j, oj, Y)) continue;
        for (int oi = -2; oi <= 2; oi++) {
          if (!couples(i, oi, X)) continue;
          const uint8_t bn = bc[m + oi + X * oj + XY * ok];
          nfree += 3 - ((bn & 1) + ((bn >> 1) & 1) + ((bn >> 2) & 1));
        }
      }
    }
    for (int r = 0; r < 3; r++) {
      nnz += ((bc[m] >> r) & 1) ? 1 : nfree;
      if (rowptr) rowptr[3 * m + r + 1] = nnz;
    }
  }
  *nnz_out = nnz;
  if (!colidx) return PTMF_OK;
  // pass 2: the stencil, one offset (9 N values) at a time; offsets in
  // ascending order give ascending neighbour nodes, so every row comes out
  // sorted
  double *S = L.galerkin ? L.S64 : nullptr;
  if (!S) {
    PTMF_TRY(ptmf::dalloc((void **)&S, sizeof(double) * 1125 * N));
    const int st = ptmf::probe_stencil(mg, level, S);
    if (st != PTMF_OK) {
      (void)hipFree(S);
      return st;
    }
  }
  std::vector<long> cur(rowptr, rowptr + 3 * N);
  std::vector<double> slab(9 * N);
  int st = PTMF_OK;
  for (int o = 0; o < 125 && st == PTMF_OK; o++) {
    const int oi = o % 5 - 2, oj = (o / 5) % 5 - 2, ok = o / 25 - 2;
    const hipError_t e = hipMemcpy(slab.data(), S + (long)o * 9 * N, sizeof(double) * 9 * N, hipMemcpyDeviceToHost);
    if (e != hipSuccess) {
      st = ptmf_set_error(PTMF_ERR_DEVICE, "mg_level_assemble: %s", hipGetErrorString(e));
      break;
    }
    for (long m = 0; m < N; m++) {
      const long i = m % X, j = (m / X) % Y, k = m / XY;
      if (!couples(i, oi, X) || !couples(j, oj, Y) || !couples(k, ok, Z)) continue;
      const long n = m + oi + X * oj + XY * ok;
      for (int r = 0; r < 3; r++) {
        const long row = 3 * m + r;
        if ((bc[m] >> r) & 1) {
          if (o == 62) {  // offset (0, 0, 0): the unit diagonal
            colidx[cur[row]] = row;
            vals[cur[row]++] = slab[(3 * r + r) * N + m];
          }
          continue;
        }
        for (int c = 0; c < 3; c++) {
          if ((bc[n] >> c) & 1) continue;
          colidx[cur[row]] = 3 * n + c;
          vals[cur[row]++] = slab[(3 * r + c) * N + m];
        }
      }
    }
  }
  if (!L.galerkin) (void)hipFree(S);
  PTMF_TRY(st);
  for (long row = 0; row < 3 * N; row++)
    PTMF_CHECK(cur[row] == rowptr[row + 1], PTMF_ERR_ARG,
               "mg_level_assemble: rowptr does not match the level's structure (row %ld)", row);
  return PTMF_OK;
}

// The six rigid-body modes of a level from its node coordinates (the near
// null space the reference attaches to its assembled levels,
// ptatin_driver_linear_ts.c:663-666): modes[k * 3N + 3n + c], k = 0, 1, 2
// translations along x, y, z; k = 3 rotation about z (-y, x, 0), k = 4 about
// x (0, -z, y), k = 5 about y (z, 0, -x).  Host array of 6 * 3N doubles, not
// normalised.
int ptmf_mg_level_rigid_body_modes(ptmf_mg mg, int level, double *modes) {
  PTMF_CHECK(mg && modes && level >= 0 && level < mg->L, PTMF_ERR_ARG, "mg_level_rigid_body_modes: bad arguments");
  ptmf::MGLevel &L = mg->lev[level];
  PTMF_CHECK(L.op->has_coords, PTMF_ERR_NOT_SETUP, "mg_level_rigid_body_modes: no coordinates");
  const long N = L.op->box.nnodes, n3 = 3 * N;
  std::vector<double> x(n3);
  PTMF_HIP(hipMemcpy(x.data(), L.op->d64.coords, sizeof(double) * n3, hipMemcpyDeviceToHost));
  std::fill(modes, modes + 6 * n3, 0.0);
  for (long n = 0; n < N; n++) {
    const double px = x[3 * n], py = x[3 * n + 1], pz = x[3 * n + 2];
    for (int c = 0; c < 3; c++) modes[c * n3 + 3 * n + c] = 1.0;
    modes[3 * n3 + 3 * n] = -py;
    modes[3 * n3 + 3 * n + 1] = px;
    modes[4 * n3 + 3 * n + 1] = -pz;
    modes[4 * n3 + 3 * n + 2] = py;
    modes[5 * n3 + 3 * n] = pz;
    modes[5 * n3 + 3 * n + 2] = -px;
  }
  return PTMF_OK;
}

int ptmf_mg_prolong_add(ptmf_mg mg, int level, const double *xc, double *yf) {
  PTMF_CHECK(mg && xc && yf && level >= 1 && level < mg->L, PTMF_ERR_ARG, "mg_prolong_add: bad level %d", level);
  PTMF_TRY(ptmf::prolong_add<double>(mg, level, xc, yf));
  PTMF_HIP(hipStreamSynchronize(mg->stream));
  return PTMF_OK;
}

int ptmf_mg_restrict(ptmf_mg mg, int level, const double *xf, double *yc) {
  PTMF_CHECK(mg && xf && yc && level >= 1 && level < mg->L, PTMF_ERR_ARG, "mg_restrict: bad level %d", level);
  PTMF_TRY(ptmf::restrict_<double>(mg, level, xf, yc));
  PTMF_HIP(hipStreamSynchronize(mg->stream));
  return PTMF_OK;
}

int ptmf_mg_smooth(ptmf_mg mg, int level, int precision, int its, int nonzero_guess, const void *b, void *x) {
  PTMF_CHECK(mg && b && x && level >= 0 && level < mg->L, PTMF_ERR_ARG, "mg_smooth: bad level %d", level);
  PTMF_CHECK(mg->is_setup, PTMF_ERR_NOT_SETUP, "mg_smooth: not set up");
  MGLevel &L = mg->lev[level];
  PTMF_CHECK(L.emax > 0, PTMF_ERR_NOT_SETUP, "mg_smooth: level %d has no eigenvalue bounds", level);
  if (precision == 64)
    PTMF_TRY(ptmf::cheb<double>(mg, L, (const double *)b, (double *)x, its, nonzero_guess != 0));
  else if (precision == 32)
    PTMF_TRY(ptmf::cheb<float>(mg, L, (const float *)b, (float *)x, its, nonzero_guess != 0));
  else
    return ptmf_set_error(PTMF_ERR_UNSUPPORTED, "mg_smooth: precision %d", precision);
  PTMF_HIP(hipStreamSynchronize(mg->stream));
  return PTMF_OK;
}

int ptmf_mg_fgmres(ptmf_mg mg, const double *b, double *x, double rtol, int maxit, int restart,
                   int *its_out, double *rnorm_out, double *history, int history_len) {
  PTMF_CHECK(mg && b && x && restart > 0 && restart < ptmf::kMaxDots && maxit > 0, PTMF_ERR_ARG,
             "mg_fgmres: bad arguments");
  PTMF_CHECK(mg->is_setup, PTMF_ERR_NOT_SETUP, "mg_fgmres: not set up");
  MGLevel &F = mg->lev[mg->L - 1];
  auto apply = [&](const double *v, double *y) { return ptmf::matmult<double>(F, v, y); };
  auto prec = [&](const double *r, double *z) { return ptmf::vcycle_fixed(mg, r, z); };
  return ptmf::fgmres_core(mg, mg->ku, F.n, apply, prec, b, x, rtol, maxit, restart, its_out, rnorm_out,
                           history, history_len);
}


// Flexible conjugate gradients preconditioned by the V-cycle (Polak-Ribiere
// beta = z_k.(r_k - r_{k-1}) / (z_{k-1}.r_{k-1}), robust to the small
// nonlinearity of an FP32 V-cycle).  Starts from the given x; converged when
// the FP64 recursive residual drops below rtol * ||b||, after which the true
// residual is recomputed (and iteration resumes if it is not below rtol).
int ptmf_mg_fcg(ptmf_mg mg, const double *b, double *x, double rtol, int maxit, int *its_out,
                double *rnorm_out, double *history, int history_len) {
  PTMF_CHECK(mg && b && x && maxit > 0, PTMF_ERR_ARG, "mg_fcg: bad arguments");
  PTMF_CHECK(mg->is_setup, PTMF_ERR_NOT_SETUP, "mg_fcg: not set up");
  ptmf::MGLevel &F = mg->lev[mg->L - 1];
  const long n = F.n;
  hipStream_t s = mg->stream;
  for (auto &p : mg->fcg)
    if (!p) PTMF_TRY(ptmf::dalloc((void **)&p, sizeof(double) * n));
  double *r = mg->fcg[0], *z = mg->fcg[1], *p = mg->fcg[2], *q = mg->fcg[3];
  double bb;
  PTMF_TRY(ptmf::dot(mg, n, b, b, &bb));
  const double bnorm = sqrt(bb);
  int its = 0, hl = 0;
  double rnorm = 0.0;
  if (bnorm == 0.0) {
    PTMF_HIP(hipMemsetAsync(x, 0, sizeof(double) * n, s));
    PTMF_HIP(hipStreamSynchronize(s));
    if (its_out) *its_out = 0;
    if (rnorm_out) *rnorm_out = 0.0;
    return PTMF_OK;
  }
  while (true) {
    // true residual
    PTMF_TRY(ptmf::matmult<double>(F, x, q));
    ptmf::k_residual<double><<<ptmf::nblocks(n), ptmf::kBS, 0, s>>>(n, b, q, r);
    PTMF_HIP(hipGetLastError());
    double rr;
    PTMF_TRY(ptmf::dot(mg, n, r, r, &rr));
    rnorm = sqrt(rr);
    PTMF_CHECK(isfinite(rnorm), PTMF_ERR_DIVERGED, "fcg: non-finite residual");
    if (history && hl < history_len) history[hl++] = rnorm;
    if (rnorm <= rtol * bnorm || its >= maxit) break;
    PTMF_TRY(ptmf_mg_vcycle(mg, r, z));
    PTMF_HIP(hipMemcpyAsync(p, z, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    double rz;
    PTMF_TRY(ptmf::dot(mg, n, r, z, &rz));
    while (its < maxit) {
      PTMF_TRY(ptmf::matmult<double>(F, p, q));
      double pq;
      PTMF_TRY(ptmf::dot(mg, n, p, q, &pq));
      PTMF_CHECK(isfinite(pq) && pq != 0.0, PTMF_ERR_DIVERGED, "fcg: breakdown (p.Ap = %g)", pq);
      const double alpha = rz / pq;
      ptmf::k_fcg_step<<<ptmf::nblocks(n), ptmf::kBS, 0, s>>>(n, alpha, p, q, x, r);
      PTMF_HIP(hipGetLastError());
      // z = M r, then r.r, z.r and z.q in one fused reduction (one host
      // synchronisation; the convergence test uses r.r).  The flexible
      // (Polak-Ribiere) beta = z.(r - r_old) / rz_old with r_old = r + alpha q
      // is -alpha z.q / rz_old: no copy of the old residual.
      PTMF_TRY(ptmf_mg_vcycle(mg, r, z));
      double d3[3];
      PTMF_TRY(ptmf::dot3(mg, n, r, z, q, d3));
      its++;
      rr = d3[0];
      rnorm = sqrt(rr);
      PTMF_CHECK(isfinite(rnorm), PTMF_ERR_DIVERGED, "fcg: non-finite residual");
      if (history && hl < history_len) history[hl++] = rnorm;
      if (rnorm <= rtol * bnorm) break;
      const double zr = d3[1], zq = d3[2];
      const double beta = -alpha * zq / rz;
      rz = zr;
      PTMF_TRY(ptmf::axpby<double>(s, n, 1.0, z, beta, p));
    }
  }
  PTMF_HIP(hipStreamSynchronize(s));
  if (its_out) *its_out = its;
  if (rnorm_out) *rnorm_out = rnorm;
  return PTMF_OK;
}


// Saddle-point set-up on the finest level: inverse Schur blocks of the
// scaled viscous mass (needs the multigrid set up for the velocity block).
int ptmf_mg_stokes_setup(ptmf_mg mg, int inner_its, double inner_rtol) {
  PTMF_CHECK(mg && inner_its >= 0 && inner_rtol >= 0.0, PTMF_ERR_ARG, "mg_stokes_setup: bad arguments");
  mg->s_inner_its = inner_its;
  mg->s_inner_rtol = inner_rtol;
  PTMF_CHECK(mg->is_setup, PTMF_ERR_NOT_SETUP, "mg_stokes_setup: multigrid not set up");
  MGLevel &F = mg->lev[mg->L - 1];
  if (!mg->sMinv) PTMF_TRY(ptmf::dalloc((void **)&mg->sMinv, sizeof(double) * 16 * F.op->box.nel));
  if (!mg->stmp) PTMF_TRY(ptmf::dalloc((void **)&mg->stmp, sizeof(double) * F.n));
  const long nup = F.n + 4 * F.op->box.nel;
  if (!mg->sx2) PTMF_TRY(ptmf::dalloc((void **)&mg->sx2, sizeof(double) * nup));
  if (!mg->sy2) PTMF_TRY(ptmf::dalloc((void **)&mg->sy2, sizeof(double) * nup));
  if (mg->opt.smoother_precision == 32) {
    if (!mg->sx2f) PTMF_TRY(ptmf::dalloc((void **)&mg->sx2f, sizeof(float) * nup));
    if (!mg->sy2f) PTMF_TRY(ptmf::dalloc((void **)&mg->sy2f, sizeof(float) * nup));
  }
  PTMF_TRY(ptmf::launch_schur_blocks(F.op, nullptr, mg->sMinv));
  PTMF_HIP(hipStreamSynchronize(mg->stream));
  mg->stokes_ready = true;
  return PTMF_OK;
}

// FGMRES on the saddle point with the block-triangular preconditioner;
// b and x are [u (3N) | p (4E)] device vectors.
int ptmf_mg_stokes_fgmres(ptmf_mg mg, const double *b, double *x, double rtol, int maxit, int restart,
                          int *its_out, double *rnorm_out, double *history, int history_len) {
  PTMF_CHECK(mg && b && x && restart > 0 && restart < ptmf::kMaxDots && maxit > 0, PTMF_ERR_ARG,
             "mg_stokes_fgmres: bad arguments");
  PTMF_CHECK(mg->stokes_ready && mg->is_setup, PTMF_ERR_NOT_SETUP, "mg_stokes_fgmres: not set up");
  MGLevel &F = mg->lev[mg->L - 1];
  const long n = F.n + 4 * F.op->box.nel;
  auto apply = [&](const double *v, double *y) { return ptmf::stokes_apply(mg, v, y); };
  auto prec = [&](const double *r, double *z) { return ptmf::stokes_prec(mg, r, z); };
  return ptmf::fgmres_core(mg, mg->ks, n, apply, prec, b, x, rtol, maxit, restart, its_out, rnorm_out,
                           history, history_len);
}

// y = K x for the saddle point ([u | p] device vectors)
// Schur approximation of the saddle-point preconditioner: 0 the scaled
// viscous mass (default, set by ptmf_mg_stokes_setup), 1 element blocks of
// B^T D^-1 B with D the finest level's A11 diagonal
int ptmf_mg_stokes_set_schur(ptmf_mg mg, int type) {
  PTMF_CHECK(mg && mg->sMinv && mg->is_setup, PTMF_ERR_NOT_SETUP, "mg_stokes_set_schur: needs ptmf_mg_stokes_setup");
  PTMF_CHECK(type == 0 || type == 1, PTMF_ERR_UNSUPPORTED, "mg_stokes_set_schur: type %d", type);
  ptmf::MGLevel &F = mg->lev[mg->L - 1];
  if (type == 0) PTMF_TRY(ptmf::launch_schur_blocks(F.op, nullptr, mg->sMinv));
  else {
    // the plain diagonal (the smoother's may be the l1 row sums), cleared
    // before the element sums; stmp is free outside a preconditioner
    // application
    PTMF_TRY(ptmf_a11_diagonal(F.op, mg->stmp));
    int *dflag = nullptr, flag = 0;
    PTMF_TRY(ptmf::dalloc((void **)&dflag, sizeof(int)));
    int st = PTMF_OK;
    if (hipMemsetAsync(dflag, 0, sizeof(int), mg->stream) != hipSuccess)
      st = ptmf_set_error(PTMF_ERR_DEVICE, "mg_stokes_set_schur: memset");
    if (st == PTMF_OK) st = ptmf::launch_schur_bdb(F.op, mg->stmp, mg->sMinv, dflag);
    if (st == PTMF_OK && hipMemcpyAsync(&flag, dflag, sizeof(int), hipMemcpyDeviceToHost, mg->stream) != hipSuccess)
      st = ptmf_set_error(PTMF_ERR_DEVICE, "mg_stokes_set_schur: flag copy");
    if (st == PTMF_OK && hipStreamSynchronize(mg->stream) != hipSuccess)
      st = ptmf_set_error(PTMF_ERR_DEVICE, "mg_stokes_set_schur: synchronize");
    (void)hipFree(dflag);
    if (st != PTMF_OK) return st;
    if (flag) {
      // keep a usable preconditioner: back to the viscous-mass blocks
      PTMF_TRY(ptmf::launch_schur_blocks(F.op, nullptr, mg->sMinv));
      PTMF_HIP(hipStreamSynchronize(mg->stream));
      return ptmf_set_error(PTMF_ERR_ARG,
                            "mg_stokes_set_schur: a B^T D^-1 B element block is singular (an element with fewer "
                            "than 4 independent free velocity rows); the viscous-mass blocks stay in use");
    }
  }
  PTMF_HIP(hipStreamSynchronize(mg->stream));
  return PTMF_OK;
}

int ptmf_mg_stokes_apply(ptmf_mg mg, const double *x, double *y) {
  PTMF_CHECK(mg && x && y, PTMF_ERR_ARG, "mg_stokes_apply: null");
  PTMF_CHECK(mg->is_setup, PTMF_ERR_NOT_SETUP, "mg_stokes_apply: not set up");
  PTMF_TRY(ptmf::stokes_apply(mg, x, y));
  PTMF_HIP(hipStreamSynchronize(mg->stream));
  return PTMF_OK;
}


// Material-point viscosity projected to the quadrature points of op's box:
// Q1-weighted nodal field from the points, then clamped corner interpolation
// (element-major eta, 27E).  Device arrays; elem holds 64-bit element indices.
int ptmf_op_project_points(ptmf_op op, long npts, const long *elem, const double *xi, const double *val,
                           double *eta) {
  PTMF_CHECK(op && eta && (npts == 0 || (elem && xi && val)), PTMF_ERR_ARG, "project_points: null");
  PTMF_CHECK(npts >= 0, PTMF_ERR_ARG, "project_points: %ld points", npts);
  const ptmf::Box &b = op->box;
  double *A = nullptr, *B = nullptr;
  PTMF_TRY(ptmf::dalloc((void **)&A, sizeof(double) * 2 * b.nnodes));
  B = A + b.nnodes;
  hipStream_t s = op->stream;
  int st = PTMF_OK;
  if (hipMemsetAsync(A, 0, sizeof(double) * 2 * b.nnodes, s) != hipSuccess)
    st = ptmf_set_error(PTMF_ERR_DEVICE, "project_points: memset");
  if (st == PTMF_OK && npts > 0) {
    ptmf::k_q1_points_accum<<<ptmf::nblocks(npts), ptmf::kBS, 0, s>>>(npts, b.mx, b.my, elem, xi, val, A, B);
    if (hipGetLastError() != hipSuccess) st = ptmf_set_error(PTMF_ERR_DEVICE, "project_points: launch");
  }
  if (st == PTMF_OK) {
    ptmf::k_divide<<<ptmf::nblocks(b.nnodes), ptmf::kBS, 0, s>>>(b.nnodes, A, B, A);
    ptmf::k_q1_to_qp<<<ptmf::nblocks(27 * b.nel), ptmf::kBS, 0, s>>>(b.mx, b.my, b.mz, A, eta);
    if (hipGetLastError() != hipSuccess) st = ptmf_set_error(PTMF_ERR_DEVICE, "project_points: launch");
  }
  if (hipStreamSynchronize(s) != hipSuccess && st == PTMF_OK)
    st = ptmf_set_error(PTMF_ERR_DEVICE, "project_points: synchronize");
  (void)hipFree(A);
  return st;
}

}  // extern "C"

// ===========================================================================
// Batched multigrid: many independent boxes (each a set-up ptmf_mg with its
// own sizes, coefficients, eigenvalue bounds and dense coarse inverse) run
// the V-cycle and flexible CG in lockstep.  Every stage is one launch over
// all boxes: the level MatMults through the batched tile kernels, the vector
// stages on a (chunk, box) grid over per-box pointer tables, the transfers
// with per-box lattice sizes and the coarse solve as one dense product per
// box.  The V-cycle is captured once as a HIP graph.
// ===========================================================================
struct ptmf_bmg_s;

namespace ptmf {
namespace {

// per-level vector roles of the pointer tables (P64 and P32 take three slots)
enum BRole {
  kB64 = 0, kX64 = 1, kR64 = 2, kP64a = 3, kD64 = 6,
  kB32 = 7, kX32 = 8, kR32 = 9, kP32a = 10, kD32 = 13, kNRoles = 14
};
constexpr int kSegBlocks = 32;  // blocks per box of the vector stages
constexpr int kMaxChebIts = 64;

template <typename T>
__global__ void kb_fill(const int *__restrict__ act, const long *__restrict__ n, T *const *y, T v) {
  const int b = blockIdx.y;
  if (act && !act[b]) return;
  T *yb = y[b];
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n[b]; i += (long)gridDim.x * blockDim.x) yb[i] = v;
}

template <typename Ts, typename Td>
__global__ void kb_cast(const int *__restrict__ act, const long *__restrict__ n, const Ts *const *x, Td *const *y) {
  const int b = blockIdx.y;
  if (act && !act[b]) return;
  const Ts *xb = x[b];
  Td *yb = y[b];
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n[b]; i += (long)gridDim.x * blockDim.x)
    yb[i] = (Td)xb[i];
}

template <typename T>
__global__ void kb_residual(const int *__restrict__ act, const long *__restrict__ n, const T *const *bb, const T *const *Ax, T *const *r) {
  const int b = blockIdx.y;
  if (act && !act[b]) return;
  const T *bv = bb[b], *av = Ax[b];
  T *rv = r[b];
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n[b]; i += (long)gridDim.x * blockDim.x)
    rv[i] = bv[i] - av[i];
}

// first Chebyshev step, per-box scale: p = x + scale dinv (b - Ax) (x, Ax may be absent)
template <typename T>
__global__ void kb_cheb_first(const int *__restrict__ act, const long *__restrict__ n, const T *const *bb, const T *const *Ax, const T *const *x,
                              const T *const *dinv, const double *__restrict__ scale, T *const *p) {
  const int b = blockIdx.y;
  if (act && !act[b]) return;
  const T *bv = bb[b], *av = Ax ? Ax[b] : nullptr, *xv = x ? x[b] : nullptr, *dv = dinv[b];
  T *pv = p[b];
  const T sc = (T)scale[b];
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n[b]; i += (long)gridDim.x * blockDim.x) {
    const T r = av ? bv[i] - av[i] : bv[i];
    const T x0 = xv ? xv[i] : T(0);
    pv[i] = x0 + sc * (r * dv[i]);
  }
}

// p_kp1 = c0 dinv (b - A p_k) + c1 p_km1 + c2 p_k with per-box (c0, c1, c2)
template <typename T>
__global__ void kb_cheb_step(const int *__restrict__ act, const long *__restrict__ n, const T *const *bb, const T *const *Apk, const T *const *dinv,
                             const T *const *pkm1, const T *const *pk, const double *__restrict__ coef,
                             T *const *pkp1) {
  const int b = blockIdx.y;
  if (act && !act[b]) return;
  const T *bv = bb[b], *av = Apk[b], *dv = dinv[b], *pm = pkm1[b], *pc = pk[b];
  T *pn = pkp1[b];
  const T c0 = (T)coef[3 * b], c1 = (T)coef[3 * b + 1], c2 = (T)coef[3 * b + 2];
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n[b]; i += (long)gridDim.x * blockDim.x)
    pn[i] = c0 * ((bv[i] - av[i]) * dv[i]) + c1 * pm[i] + c2 * pc[i];
}

template <typename T>
__global__ void kb_bc_set(const int *__restrict__ act, const long *__restrict__ nbc, const long *const *idx, T v, T *const *y) {
  const int b = blockIdx.y;
  if (act && !act[b]) return;
  const long *ib = idx[b];
  T *yb = y[b];
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < nbc[b]; i += (long)gridDim.x * blockDim.x)
    yb[ib[i]] = v;
}

// per-box lattice sizes: dims[6 b + (0..5)] = Xc, Yc, Zc, Xf, Yf, Zf
template <typename T>
__global__ void kb_restrict(const int *__restrict__ act, const long *__restrict__ dims, const T *const *xf, T *const *yc) {
  const int b = blockIdx.y;
  if (act && !act[b]) return;
  const long *d = dims + 6 * b;
  const long Xc = d[0], Yc = d[1], Zc = d[2], Xf = d[3], Yf = d[4], Zf = d[5];
  const T *xv = xf[b];
  T *yv = yc[b];
  for (long g = blockIdx.x * (long)blockDim.x + threadIdx.x; g < Xc * Yc * Zc; g += (long)gridDim.x * blockDim.x) {
    const long i = g % Xc, j = (g / Xc) % Yc, k = g / (Xc * Yc);
    T s0 = 0, s1 = 0, s2 = 0;
    for (int dk = -1; dk <= 1; dk++) {
      const long K = 2 * k + dk;
      if (K < 0 || K >= Zf) continue;
      const T wk = dk == 0 ? T(1) : T(0.5);
      for (int dj = -1; dj <= 1; dj++) {
        const long J = 2 * j + dj;
        if (J < 0 || J >= Yf) continue;
        const T wj = dj == 0 ? T(1) : T(0.5);
        for (int di = -1; di <= 1; di++) {
          const long I = 2 * i + di;
          if (I < 0 || I >= Xf) continue;
          const T w = (di == 0 ? T(1) : T(0.5)) * wj * wk;
          const long f = I + Xf * (J + Yf * K);
          s0 += w * xv[3 * f];
          s1 += w * xv[3 * f + 1];
          s2 += w * xv[3 * f + 2];
        }
      }
    }
    yv[3 * g] = s0;
    yv[3 * g + 1] = s1;
    yv[3 * g + 2] = s2;
  }
}

template <typename T>
__global__ void kb_prolong_add(const int *__restrict__ act, const long *__restrict__ dims, const T *const *xc, T *const *yf) {
  const int b = blockIdx.y;
  if (act && !act[b]) return;
  const long *d = dims + 6 * b;
  const long Xc = d[0], Yc = d[1], Xf = d[3], Yf = d[4], Zf = d[5];
  const T *xv = xc[b];
  T *yv = yf[b];
  for (long f = blockIdx.x * (long)blockDim.x + threadIdx.x; f < Xf * Yf * Zf; f += (long)gridDim.x * blockDim.x) {
    const long I = f % Xf, J = (f / Xf) % Yf, K = f / (Xf * Yf);
    const long ci0 = I >> 1, cj0 = J >> 1, ck0 = K >> 1;
    const int ni = (I & 1) ? 2 : 1, nj = (J & 1) ? 2 : 1, nk = (K & 1) ? 2 : 1;
    const T wi = ni == 2 ? T(0.5) : T(1), wj = nj == 2 ? T(0.5) : T(1), wk = nk == 2 ? T(0.5) : T(1);
    T s0 = 0, s1 = 0, s2 = 0;
    for (int c = 0; c < nk; c++)
      for (int bb = 0; bb < nj; bb++)
        for (int a = 0; a < ni; a++) {
          const long g = (ci0 + a) + Xc * ((cj0 + bb) + Yc * (ck0 + c));
          const T w = wi * wj * wk;
          s0 += w * xv[3 * g];
          s1 += w * xv[3 * g + 1];
          s2 += w * xv[3 * g + 2];
        }
    yv[3 * f] += s0;
    yv[3 * f + 1] += s1;
    yv[3 * f + 2] += s2;
  }
}

// y = A x per box with the full (symmetrised) inverse: consecutive threads
// read consecutive rows of a column
__global__ void kb_symv(const int *__restrict__ act, const long *__restrict__ n, const double *const *A, const double *const *x, double *const *y) {
  const int b = blockIdx.y;
  if (act && !act[b]) return;
  const long nb = n[b];
  const double *Ab = A[b], *xb = x[b];
  double *yb = y[b];
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < nb; i += (long)gridDim.x * blockDim.x) {
    double s = 0.0;
    for (long j = 0; j < nb; j++) s += Ab[i + j * nb] * xb[j];
    yb[i] = s;
  }
}

// per-box partial sums of up to three dot products <x_k, y_k>: part[(b * kSegBlocks + block) * 3 + k]
__global__ void kb_dots(const long *__restrict__ n, int nd, const double *const *x0, const double *const *y0,
                        const double *const *x1, const double *const *y1, const double *const *x2,
                        const double *const *y2, double *__restrict__ part) {
  __shared__ double red[3][kBS / 64];
  const int b = blockIdx.y;
  const double *a0 = x0[b], *b0 = y0[b];
  const double *a1 = nd > 1 ? x1[b] : nullptr, *b1 = nd > 1 ? y1[b] : nullptr;
  const double *a2 = nd > 2 ? x2[b] : nullptr, *b2 = nd > 2 ? y2[b] : nullptr;
  double s[3] = {0, 0, 0};
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n[b]; i += (long)gridDim.x * blockDim.x) {
    s[0] += a0[i] * b0[i];
    if (nd > 1) s[1] += a1[i] * b1[i];
    if (nd > 2) s[2] += a2[i] * b2[i];
  }
  for (int k = 0; k < 3; k++) {
    double v = s[k];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    if ((threadIdx.x & 63) == 0) red[k][threadIdx.x >> 6] = v;
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    double v = 0.0;
    for (int w = 0; w < kBS / 64; w++) v += red[threadIdx.x][w];
    part[((long)b * gridDim.x + blockIdx.x) * 3 + threadIdx.x] = v;
  }
}

// x += a p, rold = r, r -= a q (per-box a)
__global__ void kb_fcg_update(const long *__restrict__ n, const double *__restrict__ alpha, const double *const *p,
                              const double *const *q, double *const *x, double *const *r, double *const *rold) {
  const int b = blockIdx.y;
  const double a = alpha[b];
  const double *pv = p[b], *qv = q[b];
  double *xv = x[b], *rv = r[b], *ov = rold[b];
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n[b]; i += (long)gridDim.x * blockDim.x) {
    xv[i] += a * pv[i];
    ov[i] = rv[i];
    rv[i] -= a * qv[i];
  }
}

// p = z + beta p where active, per-box beta
__global__ void kb_fcg_dir(const long *__restrict__ n, const double *__restrict__ beta, const int *__restrict__ active,
                           const double *const *z, double *const *p) {
  const int b = blockIdx.y;
  if (!active[b]) return;
  const double be = beta[b];
  const double *zv = z[b];
  double *pv = p[b];
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n[b]; i += (long)gridDim.x * blockDim.x)
    pv[i] = zv[i] + be * pv[i];
}

struct BLevel {
  ptmf_batch bt = nullptr;           // batched MatMult of the boxes' level operators
  long *dn = nullptr;                // per-box vector length
  long *dnbc = nullptr;              // per-box Dirichlet count
  void **tab = nullptr;              // [kNRoles][nbox] vector pointers
  const long **bcidx = nullptr;      // per-box Dirichlet lists
  long *dims = nullptr;              // per-box lattice sizes (this level coarse, next finer fine), 6 per box
  double *scale = nullptr;           // per-box Chebyshev scale
  double *coef = nullptr;            // [kMaxChebIts][nbox][3] Chebyshev step coefficients
};

}  // namespace
}  // namespace ptmf

struct ptmf_bmg_s {
  std::vector<ptmf_mg> mgs;
  int L = 0, nbox = 0, prec = 32, pre = 0, post = 0, bc_transfer = 1;
  hipStream_t stream = nullptr;
  ptmf::BLevel lev[16];
  // coarse dense inverses
  const double **dA = nullptr;
  long *dn0 = nullptr;
  // flexible CG on the finest level: r, p, q, rold per box (one allocation) and their tables
  double *fbuf = nullptr;
  void **ftab = nullptr;  // [7][nbox]: b, x, r, p, q, rold, z
  double *part = nullptr, *hpart = nullptr;
  double *dscal = nullptr;  // [2][nbox]: alpha, beta
  int *dactive = nullptr;
  // captured V-cycle (finest B64 -> finest X64)
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int vcalls = 0;
};

namespace ptmf {
namespace {

inline dim3 seg_grid(int nbox) { return dim3(kSegBlocks, (unsigned)nbox); }

template <typename T>
T *const *btab(const BLevel &l, int role, int nbox) {
  return (T *const *)(l.tab + (long)role * nbox);
}

int bmatmult(ptmf_bmg_s *m, int l, int prec, int src, int dst) {
  const BLevel &B = m->lev[l];
  return batch_apply(B.bt, prec, (const void *const *)(B.tab + (long)src * m->nbox),
                     (void *const *)(B.tab + (long)dst * m->nbox), m->stream, m->dactive);
}

template <typename T>
int bcheb(ptmf_bmg_s *m, int l, int rb, int rx, int its, bool nonzero) {
  const BLevel &B = m->lev[l];
  const int nb = m->nbox, w = sizeof(T) == 8 ? 64 : 32;
  const int rp = sizeof(T) == 8 ? kP64a : kP32a, rr = sizeof(T) == 8 ? kR64 : kR32, rd = sizeof(T) == 8 ? kD64 : kD32;
  int km1 = 0, k = 1, kp1 = 2;
  T *const *P[3] = {btab<T>(B, rp, nb), btab<T>(B, rp + 1, nb), btab<T>(B, rp + 2, nb)};
  const int Prole[3] = {rp, rp + 1, rp + 2};
  hipStream_t s = m->stream;
  if (nonzero) {
    kb_cast<T, T><<<seg_grid(nb), kBS, 0, s>>>(m->dactive, B.dn, btab<T>(B, rx, nb), P[km1]);
    PTMF_TRY(bmatmult(m, l, w, rx, rr));
    kb_cheb_first<T><<<seg_grid(nb), kBS, 0, s>>>(m->dactive, B.dn, btab<T>(B, rb, nb), btab<T>(B, rr, nb), btab<T>(B, rx, nb),
                                                  btab<T>(B, rd, nb), B.scale, P[k]);
  } else {
    kb_fill<T><<<seg_grid(nb), kBS, 0, s>>>(m->dactive, B.dn, P[km1], T(0));
    kb_cheb_first<T><<<seg_grid(nb), kBS, 0, s>>>(m->dactive, B.dn, btab<T>(B, rb, nb), nullptr, nullptr, btab<T>(B, rd, nb),
                                                  B.scale, P[k]);
  }
  PTMF_HIP(hipGetLastError());
  for (int i = 0; i < its; i++) {
    PTMF_TRY(bmatmult(m, l, w, Prole[k], rr));
    kb_cheb_step<T><<<seg_grid(nb), kBS, 0, s>>>(m->dactive, B.dn, btab<T>(B, rb, nb), btab<T>(B, rr, nb), btab<T>(B, rd, nb),
                                                 P[km1], P[k], B.coef + (long)i * 3 * nb, P[kp1]);
    PTMF_HIP(hipGetLastError());
    const int t = km1;
    km1 = k; k = kp1; kp1 = t;
  }
  kb_cast<T, T><<<seg_grid(nb), kBS, 0, s>>>(m->dactive, B.dn, P[k], btab<T>(B, rx, nb));
  PTMF_HIP(hipGetLastError());
  return PTMF_OK;
}

// V-cycle from level l (right-hand side role rb, result role rx) down to the
// dense coarse solve on level 0; zero initial guess
template <typename T>
int bvcycle(ptmf_bmg_s *m, int l, int rb, int rx) {
  const int nb = m->nbox, w = sizeof(T) == 8 ? 64 : 32;
  const BLevel &B = m->lev[l];
  hipStream_t s = m->stream;
  if (l == 0) {
    if (sizeof(T) == 4) kb_cast<float, double><<<seg_grid(nb), kBS, 0, s>>>(m->dactive, B.dn, btab<float>(B, rb, nb),
                                                                           btab<double>(B, kB64, nb));
    kb_symv<<<seg_grid(nb), kBS, 0, s>>>(m->dactive, m->dn0, m->dA, btab<double>(B, sizeof(T) == 4 ? kB64 : rb, nb),
                                         btab<double>(B, sizeof(T) == 4 ? kX64 : rx, nb));
    if (sizeof(T) == 4) kb_cast<double, float><<<seg_grid(nb), kBS, 0, s>>>(m->dactive, B.dn, btab<double>(B, kX64, nb),
                                                                           btab<float>(B, rx, nb));
    PTMF_HIP(hipGetLastError());
    return PTMF_OK;
  }
  const BLevel &C = m->lev[l - 1];
  const int rr = sizeof(T) == 8 ? kR64 : kR32, cb = sizeof(T) == 8 ? kB64 : kB32, cx = sizeof(T) == 8 ? kX64 : kX32;
  PTMF_TRY(bcheb<T>(m, l, rb, rx, m->pre, false));
  PTMF_TRY(bmatmult(m, l, w, rx, rr));
  kb_residual<T><<<seg_grid(nb), kBS, 0, s>>>(m->dactive, B.dn, btab<T>(B, rb, nb), btab<T>(B, rr, nb), btab<T>(B, rr, nb));
  kb_restrict<T><<<seg_grid(nb), kBS, 0, s>>>(m->dactive, B.dims, btab<T>(B, rr, nb), btab<T>(C, cb, nb));
  if (m->bc_transfer) kb_bc_set<T><<<seg_grid(nb), kBS, 0, s>>>(m->dactive, C.dnbc, C.bcidx, T(0), btab<T>(C, cb, nb));
  PTMF_HIP(hipGetLastError());
  PTMF_TRY(bvcycle<T>(m, l - 1, cb, cx));
  kb_prolong_add<T><<<seg_grid(nb), kBS, 0, s>>>(m->dactive, B.dims, btab<T>(C, cx, nb), btab<T>(B, rx, nb));
  PTMF_HIP(hipGetLastError());
  return bcheb<T>(m, l, rb, rx, m->post, true);
}

// finest B64 -> finest X64
int bvcycle_body(ptmf_bmg_s *m) {
  const int lf = m->L - 1, nb = m->nbox;
  const BLevel &F = m->lev[lf];
  if (m->prec == 64) return bvcycle<double>(m, lf, kB64, kX64);
  kb_cast<double, float><<<seg_grid(nb), kBS, 0, m->stream>>>(m->dactive, F.dn, btab<double>(F, kB64, nb), btab<float>(F, kB32, nb));
  PTMF_TRY(bvcycle<float>(m, lf, kB32, kX32));
  kb_cast<float, double><<<seg_grid(nb), kBS, 0, m->stream>>>(m->dactive, F.dn, btab<float>(F, kX32, nb), btab<double>(F, kX64, nb));
  PTMF_HIP(hipGetLastError());
  return PTMF_OK;
}

int bvcycle_run(ptmf_bmg_s *m) {
  if (m->vcalls++ == 0) return bvcycle_body(m);
  if (!m->exec) {
    PTMF_HIP(hipStreamSynchronize(m->stream));
    PTMF_HIP(hipStreamBeginCapture(m->stream, hipStreamCaptureModeThreadLocal));
    const int st = bvcycle_body(m);
    hipGraph_t g = nullptr;
    const hipError_t ce = hipStreamEndCapture(m->stream, &g);
    PTMF_TRY(st);
    PTMF_CHECK(ce == hipSuccess && g, PTMF_ERR_DEVICE, "bmg: graph capture failed: %s", hipGetErrorString(ce));
    hipGraphExec_t ex = nullptr;
    const hipError_t ie = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    if (ie != hipSuccess) {
      (void)hipGraphDestroy(g);
      return ptmf_set_error(PTMF_ERR_DEVICE, "bmg: graph instantiation failed: %s", hipGetErrorString(ie));
    }
    m->graph = g;
    m->exec = ex;
  }
  PTMF_HIP(hipGraphLaunch(m->exec, m->stream));
  return PTMF_OK;
}

// per-box sums of the partial dot products (one host synchronisation)
int bdots(ptmf_bmg_s *m, int nd, const void *const *x0, const void *const *y0, const void *const *x1,
          const void *const *y1, const void *const *x2, const void *const *y2, std::vector<double> &out) {
  const int nb = m->nbox;
  const BLevel &F = m->lev[m->L - 1];
  kb_dots<<<seg_grid(nb), kBS, 0, m->stream>>>(F.dn, nd, (const double *const *)x0, (const double *const *)y0,
                                                (const double *const *)x1, (const double *const *)y1,
                                                (const double *const *)x2, (const double *const *)y2, m->part);
  PTMF_HIP(hipGetLastError());
  const size_t np = (size_t)nb * kSegBlocks * 3;
  PTMF_HIP(hipMemcpyAsync(m->hpart, m->part, sizeof(double) * np, hipMemcpyDeviceToHost, m->stream));
  PTMF_HIP(hipStreamSynchronize(m->stream));
  out.assign((size_t)nb * 3, 0.0);
  for (int b = 0; b < nb; b++)
    for (int c = 0; c < kSegBlocks; c++)
      for (int k = 0; k < 3; k++) out[3 * b + k] += m->hpart[((size_t)b * kSegBlocks + c) * 3 + k];
  return PTMF_OK;
}

}  // namespace
}  // namespace ptmf

extern "C" {

int ptmf_bmg_destroy(ptmf_bmg *pm) {
  if (!pm || !*pm) return PTMF_OK;
  ptmf_bmg_s *m = *pm;
  if (m->stream) (void)hipStreamSynchronize(m->stream);
  if (m->exec) (void)hipGraphExecDestroy(m->exec);
  if (m->graph) (void)hipGraphDestroy(m->graph);
  for (int l = 0; l < m->L; l++) {
    ptmf::BLevel &B = m->lev[l];
    if (B.bt) (void)ptmf_batch_destroy(&B.bt);
    void *ps[] = {B.dn, B.dnbc, B.tab, B.bcidx, B.dims, B.scale, B.coef};
    for (void *p : ps) (void)hipFree(p);
  }
  void *ps[] = {m->dA, m->dn0, m->fbuf, m->ftab, m->part, m->dscal, m->dactive};
  for (void *p : ps) (void)hipFree(p);
  if (m->hpart) (void)hipHostFree(m->hpart);
  if (m->stream) (void)hipStreamDestroy(m->stream);
  delete m;
  *pm = nullptr;
  return PTMF_OK;
}

int ptmf_bmg_create(const ptmf_mg *mgs, int nbox, ptmf_bmg *out) {
  PTMF_CHECK(mgs && out && nbox > 0 && nbox <= 65535, PTMF_ERR_ARG, "bmg_create: bad arguments");
  *out = nullptr;
  const ptmf_mg m0 = mgs[0];
  PTMF_CHECK(m0, PTMF_ERR_ARG, "bmg_create: null multigrid 0");
  for (int b = 0; b < nbox; b++) {
    const ptmf_mg g = mgs[b];
    PTMF_CHECK(g && g->is_setup, PTMF_ERR_NOT_SETUP, "bmg_create: multigrid %d not set up", b);
    const ptmf_mg_options &o = g->opt;
    PTMF_CHECK(g->L == m0->L && o.smoother_precision == m0->opt.smoother_precision && o.pre_its == m0->opt.pre_its &&
                   o.post_its == m0->opt.post_its && o.bc_transfer == m0->opt.bc_transfer,
               PTMF_ERR_ARG, "bmg_create: multigrid %d: levels and smoother options must match box 0", b);
    PTMF_CHECK(o.transfer == 0 && o.galerkin_levels == 0 && o.levels_mg == g->L && o.smoother_diag == 0,
               PTMF_ERR_UNSUPPORTED,
               "bmg_create: multigrid %d: needs lattice transfers, rediscretised levels, point Jacobi and the "
               "dense coarse solve", b);
    PTMF_CHECK(o.pre_its <= ptmf::kMaxChebIts && o.post_its <= ptmf::kMaxChebIts && g->L >= 2, PTMF_ERR_UNSUPPORTED,
               "bmg_create: at most %d smoothing steps and at least 2 levels", ptmf::kMaxChebIts);
    PTMF_HIP(hipStreamSynchronize(g->stream));
  }
  ptmf_bmg_s *m = new ptmf_bmg_s;
  m->mgs.assign(mgs, mgs + nbox);
  m->L = m0->L;
  m->nbox = nbox;
  m->prec = m0->opt.smoother_precision;
  m->pre = m0->opt.pre_its;
  m->post = m0->opt.post_its;
  m->bc_transfer = m0->opt.bc_transfer;
  int st = PTMF_OK;
  auto fail = [&](int code) {
    ptmf_bmg tmp = m;
    ptmf_bmg_destroy(&tmp);
    return code;
  };
  if (hipStreamCreate(&m->stream) != hipSuccess) return fail(ptmf_set_error(PTMF_ERR_DEVICE, "bmg_create: stream"));
  auto up = [&](void **dst, const void *src, size_t bytes) {
    if (st != PTMF_OK) return;
    st = ptmf::dalloc(dst, bytes);
    if (st == PTMF_OK && hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice) != hipSuccess)
      st = ptmf_set_error(PTMF_ERR_DEVICE, "bmg_create: upload");
  };
  for (int l = 0; l < m->L && st == PTMF_OK; l++) {
    ptmf::BLevel &B = m->lev[l];
    std::vector<ptmf_op> ops(nbox);
    std::vector<long> n(nbox), nbc(nbox), dims(6 * nbox, 0);
    std::vector<const long *> bci(nbox);
    std::vector<void *> tab((size_t)ptmf::kNRoles * nbox);
    std::vector<double> scale(nbox), coef((size_t)ptmf::kMaxChebIts * nbox * 3, 0.0);
    for (int b = 0; b < nbox; b++) {
      ptmf::MGLevel &Lv = m->mgs[b]->lev[l];
      ops[b] = Lv.op;
      n[b] = Lv.n;
      nbc[b] = Lv.op->nbc;
      bci[b] = Lv.op->bcidx;
      void *roles[ptmf::kNRoles] = {Lv.b64, Lv.x64, Lv.r64, Lv.p64[0], Lv.p64[1], Lv.p64[2], Lv.dinv64,
                                    Lv.b32, Lv.x32, Lv.r32, Lv.p32[0], Lv.p32[1], Lv.p32[2], Lv.dinv32};
      for (int r = 0; r < ptmf::kNRoles; r++) tab[(size_t)r * nbox + b] = roles[r];
      if (l > 0) {
        const ptmf::Box &bf = Lv.op->box, &bc = m->mgs[b]->lev[l - 1].op->box;
        const long d6[6] = {bc.X, bc.Y, bc.Z, bf.X, bf.Y, bf.Z};
        for (int k = 0; k < 6; k++) dims[6 * b + k] = d6[k];
        // the recurrence of the single-box smoother, per box
        const double emin = Lv.emin, emax = Lv.emax;
        const double sc = 2.0 / (emax + emin), alpha = 1.0 - sc * emin;
        const double mu = 1.0 / alpha, omegaprod = 2.0 / alpha;
        scale[b] = sc;
        double cm1 = 1.0, c = mu;
        for (int i = 0; i < ptmf::kMaxChebIts; i++) {
          const double cp1 = 2.0 * mu * c - cm1;
          const double omega = omegaprod * c / cp1;
          double *cf = &coef[((size_t)i * nbox + b) * 3];
          cf[0] = omega * sc;
          cf[1] = 1.0 - omega;
          cf[2] = omega;
          cm1 = c;
          c = cp1;
        }
      }
    }
    if (ptmf_batch_create(ops.data(), nbox, &B.bt) != PTMF_OK) {
      st = PTMF_ERR_DEVICE;
      break;
    }
    st = ptmf::batch_bind(B.bt, 64);
    if (st == PTMF_OK) st = ptmf::batch_bind(B.bt, 32);
    up((void **)&B.dn, n.data(), sizeof(long) * nbox);
    up((void **)&B.dnbc, nbc.data(), sizeof(long) * nbox);
    up((void **)&B.bcidx, bci.data(), sizeof(long *) * nbox);
    up((void **)&B.tab, tab.data(), sizeof(void *) * tab.size());
    up((void **)&B.dims, dims.data(), sizeof(long) * dims.size());
    up((void **)&B.scale, scale.data(), sizeof(double) * nbox);
    up((void **)&B.coef, coef.data(), sizeof(double) * coef.size());
  }
  // coarse dense inverses
  std::vector<const double *> A(nbox);
  std::vector<long> n0(nbox);
  for (int b = 0; b < nbox; b++) {
    if (!m->mgs[b]->Adense) {
      st = ptmf_set_error(PTMF_ERR_UNSUPPORTED,
                          "bmg_create: box %d has no dense coarse inverse (the batched V-cycle needs "
                          "coarse_factor 0 or 1 and at most 32768 coarsest dofs)", b);
      break;
    }
    A[b] = m->mgs[b]->Adense;
    n0[b] = m->mgs[b]->ndense;
    if (st == PTMF_OK && (!A[b] || n0[b] != m->mgs[b]->lev[0].n))
      st = ptmf_set_error(PTMF_ERR_NOT_SETUP, "bmg_create: multigrid %d has no dense coarse inverse", b);
  }
  up((void **)&m->dA, A.data(), sizeof(double *) * nbox);
  up((void **)&m->dn0, n0.data(), sizeof(long) * nbox);
  // (each box's dense inverse is already stored full, mirrored at its set-up)
  // flexible-CG workspace
  const ptmf::BLevel &F = m->lev[m->L - 1];
  (void)F;
  long tot = 0;
  std::vector<long> off(nbox);
  for (int b = 0; b < nbox; b++) {
    off[b] = tot;
    tot += m->mgs[b]->lev[m->L - 1].n;
  }
  if (st == PTMF_OK) st = ptmf::dalloc((void **)&m->fbuf, sizeof(double) * 4 * tot);
  if (st == PTMF_OK) st = ptmf::dalloc((void **)&m->ftab, sizeof(void *) * 7 * nbox);
  if (st == PTMF_OK) st = ptmf::dalloc((void **)&m->part, sizeof(double) * nbox * ptmf::kSegBlocks * 3);
  if (st == PTMF_OK) st = ptmf::dalloc((void **)&m->dscal, sizeof(double) * 2 * nbox);
  if (st == PTMF_OK) st = ptmf::dalloc((void **)&m->dactive, sizeof(int) * nbox);
  if (st == PTMF_OK && hipHostMalloc((void **)&m->hpart, sizeof(double) * nbox * ptmf::kSegBlocks * 3) != hipSuccess)
    st = ptmf_set_error(PTMF_ERR_NOMEM, "bmg_create: pinned buffer");
  if (st == PTMF_OK) {
    // tables 2..6: r, p, q, rold in the workspace, z = the V-cycle output (finest X64)
    std::vector<void *> ft((size_t)7 * nbox, nullptr);
    for (int b = 0; b < nbox; b++) {
      for (int k = 0; k < 4; k++) ft[(size_t)(2 + k) * nbox + b] = m->fbuf + (long)k * tot + off[b];
      ft[(size_t)6 * nbox + b] = m->mgs[b]->lev[m->L - 1].x64;
    }
    if (hipMemcpy(m->ftab, ft.data(), sizeof(void *) * ft.size(), hipMemcpyHostToDevice) != hipSuccess)
      st = ptmf_set_error(PTMF_ERR_DEVICE, "bmg_create: upload");
  }
  if (st != PTMF_OK) return fail(st);
  *out = m;
  return PTMF_OK;
}

// x_b = V-cycle(b_b) for every box (host arrays of device pointers, FP64)
int ptmf_bmg_vcycle(ptmf_bmg m, const double *const *b, double *const *x) {
  PTMF_CHECK(m && b && x, PTMF_ERR_ARG, "bmg_vcycle: null");
  const int nb = m->nbox;
  const ptmf::BLevel &F = m->lev[m->L - 1];
  for (int i = 0; i < nb; i++)
    PTMF_HIP(hipMemcpyAsync(m->mgs[i]->lev[m->L - 1].b64, b[i], sizeof(double) * m->mgs[i]->lev[m->L - 1].n,
                            hipMemcpyDeviceToDevice, m->stream));
  {
    std::vector<int> all(nb, 1);
    PTMF_HIP(hipMemcpyAsync(m->dactive, all.data(), sizeof(int) * nb, hipMemcpyHostToDevice, m->stream));
    PTMF_TRY(ptmf::bvcycle_run(m));
    PTMF_HIP(hipStreamSynchronize(m->stream));
  }
  for (int i = 0; i < nb; i++)
    PTMF_HIP(hipMemcpyAsync(x[i], m->mgs[i]->lev[m->L - 1].x64, sizeof(double) * m->mgs[i]->lev[m->L - 1].n,
                            hipMemcpyDeviceToDevice, m->stream));
  (void)F;
  PTMF_HIP(hipStreamSynchronize(m->stream));
  return PTMF_OK;
}

// Flexible CG on every box in lockstep, preconditioned by the batched
// V-cycle, from the given x; box b stops updating once ||r_b|| <= rtol ||b_b||.
// its[b] and rnorm[b] (true residual at the end) are per box.
int ptmf_bmg_fcg(ptmf_bmg m, const double *const *b, double *const *x, double rtol, int maxit, int *its,
                 double *rnorm) {
  PTMF_CHECK(m && b && x && its && rnorm && maxit > 0, PTMF_ERR_ARG, "bmg_fcg: bad arguments");
  const int nb = m->nbox, lf = m->L - 1;
  const ptmf::BLevel &F = m->lev[lf];
  hipStream_t s = m->stream;
  // user vectors into tables 0 (b) and 1 (x)
  std::vector<void *> ux((size_t)2 * nb);
  for (int i = 0; i < nb; i++) {
    PTMF_CHECK(b[i] && x[i], PTMF_ERR_ARG, "bmg_fcg: vectors of box %d", i);
    ux[i] = (void *)b[i];
    ux[nb + i] = x[i];
  }
  PTMF_HIP(hipMemcpyAsync(m->ftab, ux.data(), sizeof(void *) * 2 * nb, hipMemcpyHostToDevice, s));
  void *const *T = m->ftab;
  auto t = [&](int k) { return (void *const *)(T + (long)k * nb); };
  const int tB = 0, tX = 1, tR = 2, tP = 3, tQ = 4, tO = 5, tZ = 6;
  auto matmult = [&](int src, int dst) {
    return ptmf::batch_apply(F.bt, 64, (const void *const *)t(src), (void *const *)t(dst), s);
  };
  auto to_vcycle = [&]() {  // r -> finest B64
    ptmf::kb_cast<double, double><<<ptmf::seg_grid(nb), ptmf::kBS, 0, s>>>(nullptr, 
        F.dn, (const double *const *)t(tR), ptmf::btab<double>(F, ptmf::kB64, nb));
    return hipGetLastError() == hipSuccess ? PTMF_OK : ptmf_set_error(PTMF_ERR_DEVICE, "bmg_fcg: launch");
  };
  std::vector<double> d;
  // ||b||
  PTMF_TRY(ptmf::bdots(m, 1, t(tB), t(tB), nullptr, nullptr, nullptr, nullptr, d));
  std::vector<double> bn(nb), rz(nb), alpha(nb, 0.0), beta(nb, 0.0);
  std::vector<int> act(nb, 1);
  for (int i = 0; i < nb; i++) {
    bn[i] = sqrt(d[3 * i]);
    its[i] = 0;
  }
  // r = b - A x, z = V(r), p = z
  PTMF_TRY(matmult(tX, tQ));
  ptmf::kb_residual<double><<<ptmf::seg_grid(nb), ptmf::kBS, 0, s>>>(nullptr, F.dn, (const double *const *)t(tB),
                                                                       (const double *const *)t(tQ),
                                                                       (double *const *)t(tR));
  PTMF_TRY(ptmf::bdots(m, 1, t(tR), t(tR), nullptr, nullptr, nullptr, nullptr, d));
  int nact = 0;
  for (int i = 0; i < nb; i++) {
    rnorm[i] = sqrt(d[3 * i]);
    act[i] = rnorm[i] > rtol * bn[i];
    nact += act[i];
  }
  // converged boxes sit out the V-cycle (its stage kernels skip them)
  PTMF_HIP(hipMemcpyAsync(m->dactive, act.data(), sizeof(int) * nb, hipMemcpyHostToDevice, s));
  if (nact > 0) {
    PTMF_TRY(to_vcycle());
    PTMF_TRY(ptmf::bvcycle_run(m));
    ptmf::kb_cast<double, double><<<ptmf::seg_grid(nb), ptmf::kBS, 0, s>>>(nullptr, F.dn, (const double *const *)t(tZ),
                                                                             (double *const *)t(tP));
    PTMF_TRY(ptmf::bdots(m, 1, t(tR), t(tZ), nullptr, nullptr, nullptr, nullptr, d));
    for (int i = 0; i < nb; i++) rz[i] = d[3 * i];
  }
  for (int it = 0; it < maxit && nact > 0; it++) {
    PTMF_TRY(matmult(tP, tQ));
    PTMF_TRY(ptmf::bdots(m, 1, t(tP), t(tQ), nullptr, nullptr, nullptr, nullptr, d));
    for (int i = 0; i < nb; i++) {
      alpha[i] = 0.0;
      if (!act[i]) continue;
      const double pq = d[3 * i];
      PTMF_CHECK(isfinite(pq) && pq != 0.0, PTMF_ERR_DIVERGED, "bmg_fcg: box %d breakdown (p.Ap = %g)", i, pq);
      alpha[i] = rz[i] / pq;
    }
    PTMF_HIP(hipMemcpyAsync(m->dscal, alpha.data(), sizeof(double) * nb, hipMemcpyHostToDevice, s));
    ptmf::kb_fcg_update<<<ptmf::seg_grid(nb), ptmf::kBS, 0, s>>>(
        F.dn, m->dscal, (const double *const *)t(tP), (const double *const *)t(tQ), (double *const *)t(tX),
        (double *const *)t(tR), (double *const *)t(tO));
    PTMF_TRY(to_vcycle());
    PTMF_TRY(ptmf::bvcycle_run(m));
    PTMF_TRY(ptmf::bdots(m, 3, t(tR), t(tR), t(tZ), t(tR), t(tZ), t(tO), d));
    nact = 0;
    for (int i = 0; i < nb; i++) {
      if (!act[i]) continue;
      its[i]++;
      rnorm[i] = sqrt(d[3 * i]);
      PTMF_CHECK(isfinite(rnorm[i]), PTMF_ERR_DIVERGED, "bmg_fcg: box %d: non-finite residual", i);
      if (rnorm[i] <= rtol * bn[i]) {
        act[i] = 0;
        continue;
      }
      nact++;
      beta[i] = (d[3 * i + 1] - d[3 * i + 2]) / rz[i];
      rz[i] = d[3 * i + 1];
    }
    PTMF_HIP(hipMemcpyAsync(m->dscal + nb, beta.data(), sizeof(double) * nb, hipMemcpyHostToDevice, s));
    PTMF_HIP(hipMemcpyAsync(m->dactive, act.data(), sizeof(int) * nb, hipMemcpyHostToDevice, s));
    ptmf::kb_fcg_dir<<<ptmf::seg_grid(nb), ptmf::kBS, 0, s>>>(F.dn, m->dscal + nb, m->dactive,
                                                                (const double *const *)t(tZ), (double *const *)t(tP));
    PTMF_HIP(hipGetLastError());
  }
  // true residuals
  PTMF_TRY(matmult(tX, tQ));
  ptmf::kb_residual<double><<<ptmf::seg_grid(nb), ptmf::kBS, 0, s>>>(nullptr, F.dn, (const double *const *)t(tB),
                                                                       (const double *const *)t(tQ),
                                                                       (double *const *)t(tR));
  PTMF_TRY(ptmf::bdots(m, 1, t(tR), t(tR), nullptr, nullptr, nullptr, nullptr, d));
  for (int i = 0; i < nb; i++) rnorm[i] = sqrt(d[3 * i]);
  return PTMF_OK;
}

}  // extern "C"
