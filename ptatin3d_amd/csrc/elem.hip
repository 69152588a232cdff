// elem.hip -- per-element Q2 kernels other than the A11 apply: the A11
// diagonal and the pressure-velocity couplings B (A12) and B^T (A21).
//
// Layout of a work-group: 64 lanes, 2 elements, lanes 32h + q (q < 27) own
// quadrature point q of element h, lanes 32h + n (n < 27) own node n in the
// node phase.  The element's 27 node coordinates are staged in LDS once and
// read back as broadcasts; per-qp geometry (inverse Jacobian, weights) goes
// through LDS from the qp phase to the node phase.
#include "ptmf_internal.h"

namespace ptmf {
namespace {

__constant__ double cB[3][3];
__constant__ double cD[3][3];
__constant__ double cW[3];
__constant__ int cIsInit;

constexpr int kEPB = 2;  // elements per 64-lane work-group

struct QpGeom {
  double G[3][3];  // G[k][j] = d xi_j / d x_k
  double wdet;     // w_q * detJ
  double xq[3];    // physical position of the qp
};

// Geometry at quadrature point q = (qa,qb,qc) from the element's node
// coordinates xe[n][d] (P3D_evaluate_geometry_elementQ2,
// element_utils_q2.c:400-476, with the exact 3-point Gauss rule).
// value of a compile-time table column chosen by a lane-dependent index
__device__ __forceinline__ double pick3(int i, double v0, double v1, double v2) {
  return i == 0 ? v0 : (i == 1 ? v1 : v2);
}

// The 1-D tables as one lane sees them, in registers.  A lane that owns
// quadrature point (qa, qb, qc) holds rows: B[d][r] = kB[q_d][r]; a lane that
// owns node (i, j, k) holds columns: B[d][a] = kB[a][n_d].  tab_grad(T, x, y,
// z) is then the reference gradient of the (node, point) pair whose other
// index is (x, y, z), with static register indices when x, y, z are
// compile-time (unrolled loops).
struct LaneTab {
  double B[3][3], D[3][3];
};

__device__ __forceinline__ void lane_tab_qp(int q, LaneTab &T) {
  const int qd[3] = {q % 3, (q / 3) % 3, q / 9};
#pragma unroll
  for (int d = 0; d < 3; d++)
#pragma unroll
    for (int r = 0; r < 3; r++) {
      T.B[d][r] = pick3(qd[d], kB[0][r], kB[1][r], kB[2][r]);
      T.D[d][r] = pick3(qd[d], kD[0][r], kD[1][r], kD[2][r]);
    }
}

__device__ __forceinline__ void lane_tab_node(int n, LaneTab &T) {
  const int nd[3] = {n % 3, (n / 3) % 3, n / 9};
#pragma unroll
  for (int d = 0; d < 3; d++)
#pragma unroll
    for (int a = 0; a < 3; a++) {
      T.B[d][a] = pick3(nd[d], kB[a][0], kB[a][1], kB[a][2]);
      T.D[d][a] = pick3(nd[d], kD[a][0], kD[a][1], kD[a][2]);
    }
}

__device__ __forceinline__ void tab_grad(const LaneTab &T, int x, int y, int z, double gr[3]) {
  gr[0] = T.D[0][x] * T.B[1][y] * T.B[2][z];
  gr[1] = T.B[0][x] * T.D[1][y] * T.B[2][z];
  gr[2] = T.B[0][x] * T.B[1][y] * T.D[2][z];
}

__device__ void qp_geometry(const double (*xe)[3], int q, QpGeom &g) {
  const int qa = q % 3, qb = (q / 3) % 3, qc = q / 9;
  // the lane's rows of the 1-D tables, held in registers
  double Ba[3], Da[3], Bb[3], Db[3], Bc[3], Dc[3];
#pragma unroll
  for (int r = 0; r < 3; r++) {
    Ba[r] = pick3(qa, kB[0][r], kB[1][r], kB[2][r]);
    Da[r] = pick3(qa, kD[0][r], kD[1][r], kD[2][r]);
    Bb[r] = pick3(qb, kB[0][r], kB[1][r], kB[2][r]);
    Db[r] = pick3(qb, kD[0][r], kD[1][r], kD[2][r]);
    Bc[r] = pick3(qc, kB[0][r], kB[1][r], kB[2][r]);
    Dc[r] = pick3(qc, kD[0][r], kD[1][r], kD[2][r]);
  }
  double J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  double xq[3] = {0, 0, 0};
#pragma unroll
  for (int k = 0; k < 3; k++)
#pragma unroll
    for (int j = 0; j < 3; j++)
#pragma unroll
      for (int i = 0; i < 3; i++) {
        const int n = i + 3 * j + 9 * k;
        const double Ni = Ba[i], Nj = Bb[j], Nk = Bc[k];
        const double d0 = Da[i] * Nj * Nk, d1 = Ni * Db[j] * Nk, d2 = Ni * Nj * Dc[k];
        const double N = Ni * Nj * Nk;
#pragma unroll
        for (int d = 0; d < 3; d++) {
          const double x = xe[n][d];
          J[0][d] += d0 * x;
          J[1][d] += d1 * x;
          J[2][d] += d2 * x;
          xq[d] += N * x;
        }
      }
  double A[3][3];
  A[0][0] = J[1][1] * J[2][2] - J[2][1] * J[1][2];
  A[0][1] = J[2][1] * J[0][2] - J[0][1] * J[2][2];
  A[0][2] = J[0][1] * J[1][2] - J[1][1] * J[0][2];
  A[1][0] = J[2][0] * J[1][2] - J[1][0] * J[2][2];
  A[1][1] = J[0][0] * J[2][2] - J[2][0] * J[0][2];
  A[1][2] = J[1][0] * J[0][2] - J[0][0] * J[1][2];
  A[2][0] = J[1][0] * J[2][1] - J[2][0] * J[1][1];
  A[2][1] = J[2][0] * J[0][1] - J[0][0] * J[2][1];
  A[2][2] = J[0][0] * J[1][1] - J[1][0] * J[0][1];
  const double det = J[0][0] * A[0][0] + J[0][1] * A[1][0] + J[0][2] * A[2][0];
  const double id = 1.0 / det;
  for (int r = 0; r < 3; r++)
    for (int s = 0; s < 3; s++) g.G[r][s] = A[r][s] * id;
  g.wdet = pick3(qa, kW[0], kW[1], kW[2]) * pick3(qb, kW[0], kW[1], kW[2]) * pick3(qc, kW[0], kW[1], kW[2]) * det;
  for (int d = 0; d < 3; d++) g.xq[d] = xq[d];
}

// reference-coordinate gradient of node n's basis function at qp q
__device__ __forceinline__ void grad_ref(int n, int q, double gr[3]) {
  const int i = n % 3, j = (n / 3) % 3, k = n / 9;
  const int qa = q % 3, qb = (q / 3) % 3, qc = q / 9;
  const double Ni = cB[qa][i], Nj = cB[qb][j], Nk = cB[qc][k];
  gr[0] = cD[qa][i] * Nj * Nk;
  gr[1] = Ni * cD[qb][j] * Nk;
  gr[2] = Ni * Nj * cD[qc][k];
}

__device__ __forceinline__ long elem_node(int mx, int my, long X, long XY, long e, int n) {
  const long ei = e % mx, er = e / mx, ej = er % my, ek = er / my;
  const int i = n % 3, j = (n / 3) % 3, k = n / 9;
  return (2 * ei + i) + X * (2 * ej + j) + XY * (2 * ek + k);
}

// P1disc basis relative to the element (element_utils_q2.c:183-224)
__device__ __forceinline__ void p1_basis(const double (*xe)[3], const double xq[3], double Np[4]) {
  double avg[3] = {0, 0, 0};
  for (int n = 0; n < 27; n++)
    for (int d = 0; d < 3; d++) avg[d] += xe[n][d];
  const double L[3] = {xe[14][0] - xe[12][0], xe[16][1] - xe[10][1], xe[22][2] - xe[4][2]};
  Np[0] = 1.0;
  for (int d = 0; d < 3; d++) Np[1 + d] = (xq[d] - (1.0 / 27.0) * avg[d]) / L[d];
}

// ---- diag(A11): d[3n+c] += sum_q w detJ eta (|grad N_n|^2 + (d_c N_n)^2)
// (MFStokesWrapper_diagA11, stokes_operators_mf.c:67-131 with
// MatMultMF_Stokes_MixedFEM3d_diagB11, stokes_q2p1_mf_operators_diag_def.c:120-205)
__global__ void __launch_bounds__(64)
    a11_diag_kernel(int mx, int my, int mz, long nel, long X, long XY, int color, const double *__restrict__ xc,
                    const double *__restrict__ eta, double *__restrict__ dg) {
  __shared__ double sx[kEPB][27][3];
  __shared__ double sG[kEPB][27][10];
  const int h = threadIdx.x >> 5, t = threadIdx.x & 31;
  // elements of one parity class: their sums go in with plain read-add-write
  bool live;
  const long e = class_element(mx, my, mz, color, (long)blockIdx.x * kEPB + h, live);
  if (t < 27) {
    const long node = elem_node(mx, my, X, XY, e, t);
    for (int d = 0; d < 3; d++) sx[h][t][d] = xc[3 * node + d];
  }
  __syncthreads();
  if (t < 27) {
    QpGeom g;
    qp_geometry(sx[h], t, g);
    for (int r = 0; r < 3; r++)
      for (int s = 0; s < 3; s++) sG[h][t][3 * r + s] = g.G[r][s];
    sG[h][t][9] = g.wdet * eta[(long)t * nel + e];
  }
  __syncthreads();
  if (t < 27 && live) {
    LaneTab T;
    lane_tab_node(t, T);
    double acc[3] = {0, 0, 0};
    // one quadrature plane per trip (the unrolled 9 points keep their
    // geometry loads in registers; all 27 at once would spill)
#pragma unroll 1
    for (int qc = 0; qc < 3; qc++) {
      const double bk = pick3(qc, T.B[2][0], T.B[2][1], T.B[2][2]), dk = pick3(qc, T.D[2][0], T.D[2][1], T.D[2][2]);
#pragma unroll
      for (int qb = 0; qb < 3; qb++)
#pragma unroll
        for (int qa = 0; qa < 3; qa++) {
          const double *G = sG[h][qa + 3 * qb + 9 * qc];
          const double gr[3] = {T.D[0][qa] * T.B[1][qb] * bk, T.B[0][qa] * T.D[1][qb] * bk,
                                T.B[0][qa] * T.B[1][qb] * dk};
          double gp[3];
#pragma unroll
          for (int k = 0; k < 3; k++) gp[k] = G[3 * k] * gr[0] + G[3 * k + 1] * gr[1] + G[3 * k + 2] * gr[2];
          const double s = G[9];
          const double g2 = gp[0] * gp[0] + gp[1] * gp[1] + gp[2] * gp[2];
#pragma unroll
          for (int c = 0; c < 3; c++) acc[c] += s * (g2 + gp[c] * gp[c]);
        }
    }
    const long node = elem_node(mx, my, X, XY, e, t);
    for (int c = 0; c < 3; c++) dg[3 * node + c] += acc[c];
  }
}

// ---- l1 row sums of A11 with the MatMult's boundary handling, an upper
// bound of sum_j |a_ij| formed element by element:
//   l1[3n+c] += sum_{(m,e) not Dirichlet} |A_el[(n,c),(m,e)]|,
//   A_el[(n,c),(m,e)] = sum_q w detJ eta (delta_ce grad N_n . grad N_m + d_e N_n d_c N_m)
// (the element matrix of stokes_assembly.c:219-268).  Dirichlet rows are set
// to 1 afterwards (identity rows).  One element per 128-lane work-group;
// lane t < 81 owns element row t = 3 a + c.  Setup-time work.
__global__ void __launch_bounds__(128)
    a11_l1_kernel(int mx, int my, long nel, long X, long XY, const double *__restrict__ xc,
                  const double *__restrict__ eta, const uint8_t *__restrict__ bcn,
                  double *__restrict__ l1) {
  __shared__ double sx[27][3];
  __shared__ double sG[27][10];
  __shared__ double sgp[27][27][3];  // [q][node][k]: physical gradient of N_node at qp q
  __shared__ unsigned sm[27];
  const long e = blockIdx.x;
  const int t = threadIdx.x;
  if (t < 27) {
    const long node = elem_node(mx, my, X, XY, e, t);
    for (int d = 0; d < 3; d++) sx[t][d] = xc[3 * node + d];
    sm[t] = bcn ? (unsigned)bcn[node] : 0u;
  }
  __syncthreads();
  if (t < 27) {
    QpGeom g;
    qp_geometry(sx, t, g);
    for (int r = 0; r < 3; r++)
      for (int k = 0; k < 3; k++) sG[t][3 * r + k] = g.G[r][k];
    sG[t][9] = g.wdet * eta[(long)t * nel + e];
  }
  __syncthreads();
  for (int idx = t; idx < 729; idx += 128) {
    const int q = idx / 27, n = idx - 27 * q;
    double gr[3];
    grad_ref(n, q, gr);
    const double *G = sG[q];
    for (int k = 0; k < 3; k++) sgp[q][n][k] = G[3 * k] * gr[0] + G[3 * k + 1] * gr[1] + G[3 * k + 2] * gr[2];
  }
  __syncthreads();
  if (t < 81) {
    const int a = t / 3, c = t - 3 * a;
    double acc[27][3];
#pragma unroll
    for (int b = 0; b < 27; b++) acc[b][0] = acc[b][1] = acc[b][2] = 0.0;
    for (int q = 0; q < 27; q++) {
      const double s = sG[q][9];
      const double ga0 = sgp[q][a][0], ga1 = sgp[q][a][1], ga2 = sgp[q][a][2];
#pragma unroll
      for (int b = 0; b < 27; b++) {
        const double gb0 = sgp[q][b][0], gb1 = sgp[q][b][1], gb2 = sgp[q][b][2];
        const double gbc = c == 0 ? gb0 : (c == 1 ? gb1 : gb2);
        const double dot = s * (ga0 * gb0 + ga1 * gb1 + ga2 * gb2);
        const double sgbc = s * gbc;
        acc[b][0] += ga0 * sgbc;
        acc[b][1] += ga1 * sgbc;
        acc[b][2] += ga2 * sgbc;
        // delta_ce term; the branches keep the register indices static
        if (c == 0) acc[b][0] += dot;
        else if (c == 1) acc[b][1] += dot;
        else acc[b][2] += dot;
      }
    }
    double sum = 0.0;
#pragma unroll
    for (int b = 0; b < 27; b++) {
      const unsigned mb = sm[b];
#pragma unroll
      for (int ee = 0; ee < 3; ee++)
        if (!((mb >> ee) & 1u)) sum += fabs(acc[b][ee]);
    }
    const long node = elem_node(mx, my, X, XY, e, a);
    atomicAdd(&l1[3 * node + c], sum);
  }
}

// ---- B p: yu[3n+c] += sum_q -w detJ (d_c N_n) p_h(x_q)
// (MFStokesWrapper_A12, stokes_operators_mf.c:788-856 with
// MatMultMF_Stokes_MixedFEM3d_A12, stokes_q2p1_mf_operators_def.c:423-516)
__global__ void __launch_bounds__(64)
    a12_kernel(int mx, int my, int mz, long nel, long X, long XY, int color, const double *__restrict__ xc,
               const double *__restrict__ p, double *__restrict__ yu) {
  __shared__ double sx[kEPB][27][3];
  __shared__ double sG[kEPB][27][10];
  const int h = threadIdx.x >> 5, t = threadIdx.x & 31;
  bool live;  // one parity class per launch: plain read-add-write below
  const long e = class_element(mx, my, mz, color, (long)blockIdx.x * kEPB + h, live);
  if (t < 27) {
    const long node = elem_node(mx, my, X, XY, e, t);
    for (int d = 0; d < 3; d++) sx[h][t][d] = xc[3 * node + d];
  }
  __syncthreads();
  if (t < 27) {
    QpGeom g;
    qp_geometry(sx[h], t, g);
    double Np[4];
    p1_basis(sx[h], g.xq, Np);
    const double ph = Np[0] * p[4 * e] + Np[1] * p[4 * e + 1] + Np[2] * p[4 * e + 2] +
                      Np[3] * p[4 * e + 3];
    for (int r = 0; r < 3; r++)
      for (int s = 0; s < 3; s++) sG[h][t][3 * r + s] = g.G[r][s];
    sG[h][t][9] = -g.wdet * ph;
  }
  __syncthreads();
  if (t < 27 && live) {
    LaneTab T;
    lane_tab_node(t, T);
    double acc[3] = {0, 0, 0};
#pragma unroll 1
    for (int qc = 0; qc < 3; qc++) {
      const double bk = pick3(qc, T.B[2][0], T.B[2][1], T.B[2][2]), dk = pick3(qc, T.D[2][0], T.D[2][1], T.D[2][2]);
#pragma unroll
      for (int qb = 0; qb < 3; qb++)
#pragma unroll
        for (int qa = 0; qa < 3; qa++) {
          const double *G = sG[h][qa + 3 * qb + 9 * qc];
          const double gr[3] = {T.D[0][qa] * T.B[1][qb] * bk, T.B[0][qa] * T.D[1][qb] * bk,
                                T.B[0][qa] * T.B[1][qb] * dk};
#pragma unroll
          for (int k = 0; k < 3; k++)
            acc[k] += G[9] * (G[3 * k] * gr[0] + G[3 * k + 1] * gr[1] + G[3 * k + 2] * gr[2]);
        }
    }
    const long node = elem_node(mx, my, X, XY, e, t);
    for (int c = 0; c < 3; c++) yu[3 * node + c] += acc[c];
  }
}

// ---- B^T u: yp[4e+i] = sum_q -w detJ NIp_i div u_h(x_q), u Dirichlet-masked
// (MFStokesWrapper_A21, stokes_operators_mf.c:858-933 with
// MatMultMF_Stokes_MixedFEM3d_A21, stokes_q2p1_mf_operators_def.c:517-538)
__global__ void __launch_bounds__(64)
    a21_kernel(int mx, int my, long nel, long X, long XY, const double *__restrict__ xc,
               const uint8_t *__restrict__ bcn, const double *__restrict__ u,
               double *__restrict__ yp) {
  __shared__ double sx[kEPB][27][3];
  __shared__ double su[kEPB][27][3];
  __shared__ double sr[kEPB][32][4];
  const int h = threadIdx.x >> 5, t = threadIdx.x & 31;
  long e = (long)blockIdx.x * kEPB + h;
  const bool live = e < nel;
  if (!live) e = nel - 1;
  if (t < 27) {
    const long node = elem_node(mx, my, X, XY, e, t);
    const unsigned m = bcn[node];
    for (int d = 0; d < 3; d++) {
      sx[h][t][d] = xc[3 * node + d];
      su[h][t][d] = ((m >> d) & 1u) ? 0.0 : u[3 * node + d];
    }
  }
  __syncthreads();
  double val[4] = {0, 0, 0, 0};
  if (t < 27) {
    QpGeom g;
    qp_geometry(sx[h], t, g);
    double Np[4];
    p1_basis(sx[h], g.xq, Np);
    double div = 0.0;
    LaneTab T;
    lane_tab_qp(t, T);
#pragma unroll
    for (int n = 0; n < 27; n++) {
      double gr[3];
      tab_grad(T, n % 3, (n / 3) % 3, n / 9, gr);
#pragma unroll
      for (int k = 0; k < 3; k++)
        div += su[h][n][k] * (g.G[k][0] * gr[0] + g.G[k][1] * gr[1] + g.G[k][2] * gr[2]);
    }
    for (int i = 0; i < 4; i++) val[i] = -g.wdet * Np[i] * div;
  }
  for (int i = 0; i < 4; i++) sr[h][t][i] = val[i];
  __syncthreads();
  if (t < 4 && live) {
    double s = 0.0;
    for (int q = 0; q < 27; q++) s += sr[h][q][t];
    yp[4 * e + t] = s;
  }
}

// ---- Schur complement approximation: per element the 4x4 block
// M_e[i][j] = -sum_q w detJ / eta_q NIp_i NIp_j of the scaled viscous
// pressure mass matrix (MatAssemble_StokesPC_ScaledMassMatrix,
// stokes_assembly.c:299-450); Minv_e its inverse (Gauss-Jordan with partial
// pivoting).  Either output may be null.
__global__ void __launch_bounds__(64)
    schur_blocks_kernel(int mx, int my, long nel, long X, long XY, const double *__restrict__ xc,
                        const double *__restrict__ eta, double *__restrict__ Me,
                        double *__restrict__ Minv) {
  __shared__ double sx[kEPB][27][3];
  __shared__ double sr[kEPB][32][10];
  const int h = threadIdx.x >> 5, t = threadIdx.x & 31;
  long e = (long)blockIdx.x * kEPB + h;
  const bool live = e < nel;
  if (!live) e = nel - 1;
  if (t < 27) {
    const long node = elem_node(mx, my, X, XY, e, t);
    for (int d = 0; d < 3; d++) sx[h][t][d] = xc[3 * node + d];
  }
  __syncthreads();
  double v[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  if (t < 27) {
    QpGeom g;
    qp_geometry(sx[h], t, g);
    double Np[4];
    p1_basis(sx[h], g.xq, Np);
    const double fac = -g.wdet / eta[(long)t * nel + e];
    int k = 0;
    for (int i = 0; i < 4; i++)
      for (int j = i; j < 4; j++) v[k++] = fac * Np[i] * Np[j];
  }
  for (int k = 0; k < 10; k++) sr[h][t][k] = v[k];
  __syncthreads();
  if (t == 0 && live) {
    double A[4][4];
    int k = 0;
    for (int i = 0; i < 4; i++)
      for (int j = i; j < 4; j++) {
        double s = 0.0;
        for (int q = 0; q < 27; q++) s += sr[h][q][k];
        A[i][j] = A[j][i] = s;
        k++;
      }
    if (Me)
      for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) Me[16 * e + 4 * i + j] = A[i][j];
    if (Minv) {
      double B[4][4] = {{1, 0, 0, 0}, {0, 1, 0, 0}, {0, 0, 1, 0}, {0, 0, 0, 1}};
      for (int c = 0; c < 4; c++) {
        int pr = c;
        for (int r = c + 1; r < 4; r++)
          if (fabs(A[r][c]) > fabs(A[pr][c])) pr = r;
        for (int j = 0; j < 4; j++) {
          const double ta = A[c][j], tb = B[c][j];
          A[c][j] = A[pr][j]; B[c][j] = B[pr][j];
          A[pr][j] = ta; B[pr][j] = tb;
        }
        const double ip = 1.0 / A[c][c];
        for (int j = 0; j < 4; j++) { A[c][j] *= ip; B[c][j] *= ip; }
        for (int r = 0; r < 4; r++) {
          if (r == c) continue;
          const double f = A[r][c];
          for (int j = 0; j < 4; j++) { A[r][j] -= f * A[c][j]; B[r][j] -= f * B[c][j]; }
        }
      }
      for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) Minv[16 * e + 4 * i + j] = B[i][j];
    }
  }
}

// ---- element blocks of B^T D^-1 B with D the assembled A11 diagonal:
// S_e[i][j] = sum_{n,d not Dirichlet} B_e[(n,d),i] B_e[(n,d),j] / D[3 node + d],
// B_e[(n,d),i] = -sum_q w detJ (d_d N_n) NIp_i (the element's A12 columns);
// Minv_e = (-S_e)^-1.  An element-local stand-in for the Schur complement
// -B^T A^-1 B that sees the stiffness of the neighbours through D.
__global__ void __launch_bounds__(64)
    schur_bdb_kernel(int mx, int my, long nel, long X, long XY, const double *__restrict__ xc,
                     const uint8_t *__restrict__ bcn, const double *__restrict__ dg, double *__restrict__ Minv,
                     int *__restrict__ singular) {
  __shared__ double sx[kEPB][27][3];
  __shared__ double sG[kEPB][27][14];  // G (9), -w detJ, NIp (4)
  __shared__ double sr[kEPB][32][10];
  const int h = threadIdx.x >> 5, t = threadIdx.x & 31;
  long e = (long)blockIdx.x * kEPB + h;
  const bool live = e < nel;
  if (!live) e = nel - 1;
  if (t < 27) {
    const long node = elem_node(mx, my, X, XY, e, t);
    for (int d = 0; d < 3; d++) sx[h][t][d] = xc[3 * node + d];
  }
  __syncthreads();
  if (t < 27) {
    QpGeom g;
    qp_geometry(sx[h], t, g);
    double Np[4];
    p1_basis(sx[h], g.xq, Np);
    for (int r = 0; r < 3; r++)
      for (int c = 0; c < 3; c++) sG[h][t][3 * r + c] = g.G[r][c];
    sG[h][t][9] = -g.wdet;
    for (int i = 0; i < 4; i++) sG[h][t][10 + i] = Np[i];
  }
  __syncthreads();
  double v[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  if (t < 27) {
    LaneTab T;
    lane_tab_node(t, T);
    double Bn[3][4] = {{0, 0, 0, 0}, {0, 0, 0, 0}, {0, 0, 0, 0}};
#pragma unroll 1
    for (int qc = 0; qc < 3; qc++) {
      const double bk = pick3(qc, T.B[2][0], T.B[2][1], T.B[2][2]), dk = pick3(qc, T.D[2][0], T.D[2][1], T.D[2][2]);
#pragma unroll
      for (int qb = 0; qb < 3; qb++)
#pragma unroll
        for (int qa = 0; qa < 3; qa++) {
          const double *G = sG[h][qa + 3 * qb + 9 * qc];
          const double gr[3] = {T.D[0][qa] * T.B[1][qb] * bk, T.B[0][qa] * T.D[1][qb] * bk,
                                T.B[0][qa] * T.B[1][qb] * dk};
#pragma unroll
          for (int d = 0; d < 3; d++) {
            const double gd = G[9] * (G[3 * d] * gr[0] + G[3 * d + 1] * gr[1] + G[3 * d + 2] * gr[2]);
#pragma unroll
            for (int i = 0; i < 4; i++) Bn[d][i] += gd * G[10 + i];
          }
        }
    }
    const long node = elem_node(mx, my, X, XY, e, t);
    const unsigned m = bcn ? (unsigned)bcn[node] : 0u;
    for (int d = 0; d < 3; d++) {
      if ((m >> d) & 1u) continue;
      const double id = 1.0 / dg[3 * node + d];
      int k = 0;
      for (int i = 0; i < 4; i++)
        for (int j = i; j < 4; j++) v[k++] += Bn[d][i] * Bn[d][j] * id;
    }
  }
  for (int k = 0; k < 10; k++) sr[h][t][k] = v[k];
  __syncthreads();
  if (t == 0 && live) {
    double A[4][4];
    int k = 0;
    for (int i = 0; i < 4; i++)
      for (int j = i; j < 4; j++) {
        double s = 0.0;
        for (int q = 0; q < 27; q++) s += sr[h][q][k];
        A[i][j] = A[j][i] = -s;
        k++;
      }
    double B[4][4] = {{1, 0, 0, 0}, {0, 1, 0, 0}, {0, 0, 1, 0}, {0, 0, 0, 1}};
    // S_e has rank <= the element's free velocity rows: a vanishing pivot
    // (fewer than 4 independent free rows) is reported, not inverted silently
    double amax = 0.0;
    for (int c = 0; c < 4; c++) amax = fmax(amax, fabs(A[c][c]));
    for (int c = 0; c < 4; c++) {
      int pr = c;
      for (int r = c + 1; r < 4; r++)
        if (fabs(A[r][c]) > fabs(A[pr][c])) pr = r;
      if (!(fabs(A[pr][c]) > 1e-13 * amax)) atomicOr(singular, 1);
      for (int j = 0; j < 4; j++) {
        const double ta = A[c][j], tb = B[c][j];
        A[c][j] = A[pr][j]; B[c][j] = B[pr][j];
        A[pr][j] = ta; B[pr][j] = tb;
      }
      const double ip = 1.0 / A[c][c];
      for (int j = 0; j < 4; j++) { A[c][j] *= ip; B[c][j] *= ip; }
      for (int r = 0; r < 4; r++) {
        if (r == c) continue;
        const double f = A[r][c];
        for (int j = 0; j < 4; j++) { A[r][j] -= f * A[c][j]; B[r][j] -= f * B[c][j]; }
      }
    }
    for (int i = 0; i < 4; i++)
      for (int j = 0; j < 4; j++) Minv[16 * e + 4 * i + j] = B[i][j];
  }
}

// yp[4e+i] = sum_j Minv_e[i][j] rp[4e+j] (+ beta yp when beta != 0)
__global__ void schur_apply_kernel(long nel, const double *__restrict__ Minv, const double *__restrict__ rp,
                                   double *__restrict__ yp) {
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < 4 * nel; t += (long)gridDim.x * blockDim.x) {
    const long e = t >> 2;
    const int i = (int)(t & 3);
    const double *M = Minv + 16 * e + 4 * i;
    const double *r = rp + 4 * e;
    yp[t] = M[0] * r[0] + M[1] * r[1] + M[2] * r[2] + M[3] * r[3];
  }
}

// ---- strain-rate-dependent viscosity at the quadrature points:
// eta = clip(B * max(eII, 1e-17)^(1/n - 1)), eII = sqrt(0.5 D:D), D = sym(grad u)
// (ComputeStrainRate3d / ComputeSecondInvariant3d, stokes_rheology_vp_std.c:85-148;
// the VISCOUS_ARRHENIUS_2 power law, stokes_rheology_vp_std.c:437-461, with the
// prefactor B per quadrature point).  Element-major input B and output eta.
__global__ void __launch_bounds__(64)
    powerlaw_kernel(int mx, int my, long nel, long X, long XY, const double *__restrict__ xc,
                    const double *__restrict__ u, const double *__restrict__ B, double nexp, double eta_min,
                    double eta_max, double *__restrict__ eta) {
  __shared__ double sx[kEPB][27][3];
  __shared__ double su[kEPB][27][3];
  const int h = threadIdx.x >> 5, t = threadIdx.x & 31;
  long e = (long)blockIdx.x * kEPB + h;
  const bool live = e < nel;
  if (!live) e = nel - 1;
  if (t < 27) {
    const long node = elem_node(mx, my, X, XY, e, t);
    for (int d = 0; d < 3; d++) {
      sx[h][t][d] = xc[3 * node + d];
      su[h][t][d] = u[3 * node + d];
    }
  }
  __syncthreads();
  if (t < 27 && live) {
    QpGeom g;
    qp_geometry(sx[h], t, g);
    double L[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};  // L[l][k] = d u_l / d x_k
    LaneTab T;
    lane_tab_qp(t, T);
#pragma unroll
    for (int n = 0; n < 27; n++) {
      double gr[3], gp[3];
      tab_grad(T, n % 3, (n / 3) % 3, n / 9, gr);
      for (int k = 0; k < 3; k++) gp[k] = g.G[k][0] * gr[0] + g.G[k][1] * gr[1] + g.G[k][2] * gr[2];
      for (int l = 0; l < 3; l++)
        for (int k = 0; k < 3; k++) L[l][k] += su[h][n][l] * gp[k];
    }
    double sum = 0.0;
    for (int l = 0; l < 3; l++)
      for (int k = 0; k < 3; k++) {
        const double dlk = 0.5 * (L[l][k] + L[k][l]);
        sum += dlk * dlk;
      }
    double sr = sqrt(0.5 * sum);
    if (sr < 1.0e-17) sr = 1.0e-17;
    double v = B[27 * e + t] * pow(sr, 1.0 / nexp - 1.0);
    v = fmin(fmax(v, eta_min), eta_max);
    eta[27 * e + t] = v;
  }
}

int init_constants() {
  static bool done = false;
  if (done) return PTMF_OK;
  PTMF_HIP(hipMemcpyToSymbol(HIP_SYMBOL(cB), kB, sizeof(kB)));
  PTMF_HIP(hipMemcpyToSymbol(HIP_SYMBOL(cD), kD, sizeof(kD)));
  PTMF_HIP(hipMemcpyToSymbol(HIP_SYMBOL(cW), kW, sizeof(kW)));
  done = true;
  return PTMF_OK;
}

}  // namespace

int launch_a11_diag(const ptmf_op_s *op, double *d) {
  PTMF_TRY(init_constants());
  const Box &b = op->box;
  for (int color = 0; color < 8; color++) {
    const long ncol = class_count(b, color);
    if (ncol == 0) continue;
    const long nblk = (ncol + kEPB - 1) / kEPB;
    a11_diag_kernel<<<dim3((unsigned)nblk), dim3(64), 0, op->stream>>>(
        b.mx, b.my, b.mz, b.nel, b.X, b.X * b.Y, color, op->d64.coords, op->d64.eta, d);
    PTMF_HIP(hipGetLastError());
  }
  return PTMF_OK;
}

int launch_a11_l1(const ptmf_op_s *op, double *d) {
  PTMF_TRY(init_constants());
  const Box &b = op->box;
  PTMF_CHECK(b.nel < (1L << 31), PTMF_ERR_ARG, "a11_l1: mesh too large");
  a11_l1_kernel<<<dim3((unsigned)b.nel), dim3(128), 0, op->stream>>>(
      b.mx, b.my, b.nel, b.X, b.X * b.Y, op->d64.coords, op->d64.eta, op->bcnode, d);
  PTMF_HIP(hipGetLastError());
  return PTMF_OK;
}

int launch_schur_blocks(const ptmf_op_s *op, double *Me, double *Minv) {
  PTMF_TRY(init_constants());
  const Box &b = op->box;
  const long nblk = (b.nel + kEPB - 1) / kEPB;
  schur_blocks_kernel<<<dim3((unsigned)nblk), dim3(64), 0, op->stream>>>(
      b.mx, b.my, b.nel, b.X, b.X * b.Y, op->d64.coords, op->d64.eta, Me, Minv);
  PTMF_HIP(hipGetLastError());
  return PTMF_OK;
}

int launch_schur_bdb(const ptmf_op_s *op, const double *dg, double *Minv, int *singular) {
  PTMF_TRY(init_constants());
  const Box &b = op->box;
  const long nblk = (b.nel + kEPB - 1) / kEPB;
  schur_bdb_kernel<<<dim3((unsigned)nblk), dim3(64), 0, op->stream>>>(b.mx, b.my, b.nel, b.X, b.X * b.Y,
                                                                      op->d64.coords, op->bcnode, dg, Minv,
                                                                      singular);
  PTMF_HIP(hipGetLastError());
  return PTMF_OK;
}

int launch_schur_apply(const ptmf_op_s *op, const double *Minv, const double *rp, double *yp) {
  const long n = 4 * op->box.nel;
  long nb = (n + 255) / 256;
  if (nb > 65536) nb = 65536;
  schur_apply_kernel<<<dim3((unsigned)nb), dim3(256), 0, op->stream>>>(op->box.nel, Minv, rp, yp);
  PTMF_HIP(hipGetLastError());
  return PTMF_OK;
}

int launch_powerlaw(const ptmf_op_s *op, const double *u, const double *B, double nexp, double eta_min,
                    double eta_max, double *eta) {
  PTMF_TRY(init_constants());
  const Box &b = op->box;
  const long nblk = (b.nel + kEPB - 1) / kEPB;
  powerlaw_kernel<<<dim3((unsigned)nblk), dim3(64), 0, op->stream>>>(b.mx, b.my, b.nel, b.X, b.X * b.Y,
                                                                     op->d64.coords, u, B, nexp, eta_min,
                                                                     eta_max, eta);
  PTMF_HIP(hipGetLastError());
  return PTMF_OK;
}

int launch_a12(const ptmf_op_s *op, const double *p, double *yu) {
  PTMF_TRY(init_constants());
  const Box &b = op->box;
  for (int color = 0; color < 8; color++) {
    const long ncol = class_count(b, color);
    if (ncol == 0) continue;
    const long nblk = (ncol + kEPB - 1) / kEPB;
    a12_kernel<<<dim3((unsigned)nblk), dim3(64), 0, op->stream>>>(b.mx, b.my, b.mz, b.nel, b.X, b.X * b.Y,
                                                                   color, op->d64.coords, p, yu);
    PTMF_HIP(hipGetLastError());
  }
  return PTMF_OK;
}

int launch_a21(const ptmf_op_s *op, const double *u, double *yp) {
  PTMF_TRY(init_constants());
  const Box &b = op->box;
  const long nblk = (b.nel + kEPB - 1) / kEPB;
  a21_kernel<<<dim3((unsigned)nblk), dim3(64), 0, op->stream>>>(
      b.mx, b.my, b.nel, b.X, b.X * b.Y, op->d64.coords, op->bcnode, u, yp);
  PTMF_HIP(hipGetLastError());
  return PTMF_OK;
}

}  // namespace ptmf
