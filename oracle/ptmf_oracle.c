/*
 * ptmf_oracle.c -- CPU restatement of pTatin3d's matrix-free Q2-P1disc Stokes
 * operators (TEST INFRASTRUCTURE: the correctness checker and the CPU
 * baseline; never linked into or called by the product library).
 *
 * Every function cites the reference file:line it restates.  Layouts are
 * documented in ptmf_oracle.h.  Pinning (which independent checks each
 * function is held to) is listed in tests/test_oracle.py.
 */
#include "ptmf_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define NB 4 /* element batch of the tensor path (NEV, include/stokes_operators.h:103) */

int orc_num_threads(void)
{
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

void orc_set_num_threads(int n)
{
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

/* ------------------------------------------------------------------------ */
/* structured box helpers (dmda_element_q2p1.c:368-469)                      */
/* ------------------------------------------------------------------------ */
typedef struct {
  long mx, my, mz, X, Y, Z, nel, nnodes;
} box_t;

static box_t mkbox(int mx, int my, int mz)
{
  box_t b;
  b.mx = mx; b.my = my; b.mz = mz;
  b.X = 2L * mx + 1; b.Y = 2L * my + 1; b.Z = 2L * mz + 1;
  b.nel = (long)mx * my * mz;
  b.nnodes = b.X * b.Y * b.Z;
  return b;
}

/* the 27 node indices of element e, local node n = i + 3j + 9k
 * (dmda_element_q2p1.c:405-439) */
static void elnodes(const box_t *b, long e, long nid[27])
{
  long ei = e % b->mx, ej = (e / b->mx) % b->my, ek = e / (b->mx * b->my);
  for (int k = 0; k < 3; k++)
    for (int j = 0; j < 3; j++)
      for (int i = 0; i < 3; i++)
        nid[i + 3 * j + 9 * k] = (2 * ei + i) + b->X * ((2 * ej + j) + b->Y * (2 * ek + k));
}

/* ------------------------------------------------------------------------ */
/* tensor path: stokes_operators_tensor.c                                    */
/* ------------------------------------------------------------------------ */

/* 3-point Gauss-Legendre rule on [-1,1].  The reference obtains it from
 * PetscDTGaussQuadrature(3,-1,1,...) (stokes_operators_tensor.c:194); the
 * closed form agrees with it to rounding. */
static void gauss3(double x[3], double w[3])
{
  const double s = sqrt(3.0 / 5.0);
  x[0] = -s; x[1] = 0.0; x[2] = s;
  w[0] = 5.0 / 9.0; w[1] = 8.0 / 9.0; w[2] = 5.0 / 9.0;
}

/* 1-D quadratic Lagrange basis (nodes -1,0,1) B[q][n] and its derivative
 * D[q][n] at the Gauss points (stokes_operators_tensor.c:195-202). */
static void basis_tables(double B[3][3], double D[3][3], double wq[27])
{
  double x[3], w[3];
  gauss3(x, w);
  for (int q = 0; q < 3; q++) {
    B[q][0] = 0.5 * (x[q] * x[q] - x[q]);
    B[q][1] = 1.0 - x[q] * x[q];
    B[q][2] = 0.5 * (x[q] * x[q] + x[q]);
    D[q][0] = x[q] - 0.5;
    D[q][1] = -2.0 * x[q];
    D[q][2] = x[q] + 0.5;
  }
  for (int c = 0; c < 3; c++)
    for (int b = 0; b < 3; b++)
      for (int a = 0; a < 3; a++) wq[(c * 3 + b) * 3 + a] = w[c] * w[b] * w[a];
}

/* Three successive 1-D contractions on a batch of NB elements and 3 fields
 * (stokes_operators_tensor.c:37-92):
 *   y[f][c][b][a] += T[c][k] S[b][j] R[a][i] x[f][k][j][i]   (a fastest).
 * With transpose != 0 the three 3x3 tables are applied transposed. */
static void contract3(const double R0[3][3], const double S0[3][3], const double T0[3][3],
                      int transpose, const double x[3][27][NB], double y[3][27][NB])
{
  double R[3][3], S[3][3], T[3][3];
  double s1[3][27][NB], s2[3][27][NB];
  for (int r = 0; r < 3; r++)
    for (int c = 0; c < 3; c++) {
      R[r][c] = transpose ? R0[c][r] : R0[r][c];
      S[r][c] = transpose ? S0[c][r] : S0[r][c];
      T[r][c] = transpose ? T0[c][r] : T0[r][c];
    }
  memset(s1, 0, sizeof s1);
  memset(s2, 0, sizeof s2);
  /* fastest index */
  for (int f = 0; f < 3; f++)
    for (int kj = 0; kj < 9; kj++)
      for (int i = 0; i < 3; i++)
        for (int a = 0; a < 3; a++)
          for (int e = 0; e < NB; e++) s1[f][kj * 3 + a][e] += R[a][i] * x[f][kj * 3 + i][e];
  /* middle index */
  for (int f = 0; f < 3; f++)
    for (int k = 0; k < 3; k++)
      for (int j = 0; j < 3; j++)
        for (int a = 0; a < 3; a++)
          for (int b = 0; b < 3; b++)
            for (int e = 0; e < NB; e++)
              s2[f][(k * 3 + b) * 3 + a][e] += S[b][j] * s1[f][(k * 3 + j) * 3 + a][e];
  /* slowest index */
  for (int k = 0; k < 3; k++)
    for (int f = 0; f < 3; f++)
      for (int c = 0; c < 3; c++)
        for (int ba = 0; ba < 9; ba++)
          for (int e = 0; e < NB; e++) y[f][c * 9 + ba][e] += T[c][k] * s2[f][k * 9 + ba][e];
}

/* In-place inverse of the per-qp Jacobian (stokes_operators_tensor.c:94-125).
 * On entry dx[j][k][q] = d x_k / d xi_j; on exit dx[k][j][q] = d xi_j / d x_k
 * and det[q] = det J. */
static void jacobian_invert(double dx[3][3][27][NB], double det[27][NB])
{
  for (int q = 0; q < 27; q++)
    for (int e = 0; e < NB; e++) {
      double a[3][3];
      for (int j = 0; j < 3; j++)
        for (int k = 0; k < 3; k++) a[j][k] = dx[j][k][q][e];
      double c0 = (a[1][1] * a[2][2] - a[2][1] * a[1][2]);
      double c3 = -(a[1][0] * a[2][2] - a[2][0] * a[1][2]);
      double c6 = (a[1][0] * a[2][1] - a[2][0] * a[1][1]);
      double d = a[0][0] * c0 + a[0][1] * c3 + a[0][2] * c6;
      double id = 1.0 / d;
      dx[0][0][q][e] = id * c0;
      dx[0][1][q][e] = -id * (a[0][1] * a[2][2] - a[2][1] * a[0][2]);
      dx[0][2][q][e] = id * (a[0][1] * a[1][2] - a[1][1] * a[0][2]);
      dx[1][0][q][e] = id * c3;
      dx[1][1][q][e] = id * (a[0][0] * a[2][2] - a[2][0] * a[0][2]);
      dx[1][2][q][e] = -id * (a[0][0] * a[1][2] - a[1][0] * a[0][2]);
      dx[2][0][q][e] = id * c6;
      dx[2][1][q][e] = -id * (a[0][0] * a[2][1] - a[2][0] * a[0][1]);
      dx[2][2][q][e] = id * (a[0][0] * a[1][1] - a[1][0] * a[0][1]);
      det[q][e] = d;
    }
}

/* Pointwise viscous stress at each qp (stokes_operators_tensor.c:127-179):
 * grad u -> sym -> 2 eta eps -> w detJ sigma G^T. */
static void qp_action(const double *eta[NB], double dx[3][3][27][NB], double det[27][NB],
                      const double wq[27], double du[3][3][27][NB], double dv[3][3][27][NB])
{
  for (int q = 0; q < 27; q++)
    for (int e = 0; e < NB; e++) {
      double g[3][3]; /* g[k][l] = d u_l / d x_k */
      for (int l = 0; l < 3; l++)
        for (int k = 0; k < 3; k++)
          g[k][l] = du[0][l][q][e] * dx[k][0][q][e] + du[1][l][q][e] * dx[k][1][q][e] +
                    du[2][l][q][e] * dx[k][2][q][e];
      double eps[6] = {g[0][0], g[1][1], g[2][2], 0.5 * (g[0][1] + g[1][0]),
                       0.5 * (g[0][2] + g[2][0]), 0.5 * (g[1][2] + g[2][1])};
      double s[6];
      for (int m = 0; m < 6; m++) s[m] = 2 * eta[e][q] * eps[m];
      double sig[3][3] = {{s[0], s[3], s[4]}, {s[3], s[1], s[5]}, {s[4], s[5], s[2]}};
      for (int l = 0; l < 3; l++)
        for (int k = 0; k < 3; k++)
          dv[k][l][q][e] = wq[q] * det[q][e] *
                           (sig[0][l] * dx[0][k][q][e] + sig[1][l] * dx[1][k][q][e] +
                            sig[2][l] * dx[2][k][q][e]);
    }
}

void orc_a11_tensor(int mx, int my, int mz, const double *coords, const double *eta,
                    const double *u, double *y)
{
  box_t b = mkbox(mx, my, mz);
  double B[3][3], D[3][3], wq[27];
  basis_tables(B, D, wq);
  const long nel = b.nel;

#pragma omp parallel for schedule(static)
  for (long e0 = 0; e0 < nel; e0 += NB) {
    double elx[3][27][NB], elu[3][27][NB], elv[3][27][NB];
    double dx[3][3][27][NB], det[27][NB], du[3][3][27][NB], dv[3][3][27][NB];
    const double *eq[NB];
    long nid[NB][27];
    for (int e = 0; e < NB; e++) {
      long ee = (e0 + e < nel) ? e0 + e : nel - 1; /* pad the last batch */
      elnodes(&b, ee, nid[e]);
      eq[e] = eta + 27 * ee;
      for (int n = 0; n < 27; n++)
        for (int l = 0; l < 3; l++) {
          elx[l][n][e] = coords[3 * nid[e][n] + l];
          elu[l][n][e] = u[3 * nid[e][n] + l];
        }
    }
    memset(dx, 0, sizeof dx);
    contract3(D, B, B, 0, elx, dx[0]);
    contract3(B, D, B, 0, elx, dx[1]);
    contract3(B, B, D, 0, elx, dx[2]);
    jacobian_invert(dx, det);
    memset(du, 0, sizeof du);
    contract3(D, B, B, 0, elu, du[0]);
    contract3(B, D, B, 0, elu, du[1]);
    contract3(B, B, D, 0, elu, du[2]);
    qp_action(eq, dx, det, wq, du, dv);
    memset(elv, 0, sizeof elv);
    contract3(D, B, B, 1, dv[0], elv);
    contract3(B, D, B, 1, dv[1], elv);
    contract3(B, B, D, 1, dv[2], elv);
#pragma omp critical
    {
      for (int e = 0; e < NB && e0 + e < nel; e++)
        for (int n = 0; n < 27; n++)
          for (int l = 0; l < 3; l++) y[3 * nid[e][n] + l] += elv[l][n][e];
    }
  }
}

/* ------------------------------------------------------------------------ */
/* "ref" element path: element_utils_q2.c + stokes_q2p1_mf_operators_*.c     */
/* ------------------------------------------------------------------------ */

/* P3D_prepare_elementQ2_3x3 (element_utils_q2.c:276-333): quadrature with
 * the reference's truncated constants, qp index q = i + 3j + 9k. */
static void prep_q2_3x3(double W[27], double XI[27][3], double NI[27][27], double GNI[27][3][27])
{
  const double s = 0.774596669241483, w1 = 0.555555555555556, w0 = 0.888888888888889;
  const double ox[3] = {-s, 0.0, s}, ow[3] = {w1, w0, w1};
  int cnt = 0;
  for (int k = 0; k < 3; k++)
    for (int j = 0; j < 3; j++)
      for (int i = 0; i < 3; i++) {
        XI[cnt][0] = ox[i]; XI[cnt][1] = ox[j]; XI[cnt][2] = ox[k];
        W[cnt] = ow[i] * ow[j] * ow[k];
        cnt++;
      }
  for (int p = 0; p < 27; p++) {
    double N1[3][3], G1[3][3];
    for (int d = 0; d < 3; d++) {
      double t = XI[p][d];
      N1[d][0] = 0.5 * t * (t - 1.0);
      N1[d][1] = (1.0 + t) * (1.0 - t);
      N1[d][2] = 0.5 * (1.0 + t) * t;
      G1[d][0] = t - 0.5;
      G1[d][1] = -2.0 * t;
      G1[d][2] = 0.5 * (1.0 + 2.0 * t);
    }
    int n = 0;
    for (int k = 0; k < 3; k++)
      for (int j = 0; j < 3; j++)
        for (int i = 0; i < 3; i++) {
          NI[p][n] = N1[0][i] * N1[1][j] * N1[2][k];
          GNI[p][0][n] = G1[0][i] * N1[1][j] * N1[2][k];
          GNI[p][1][n] = N1[0][i] * G1[1][j] * N1[2][k];
          GNI[p][2][n] = N1[0][i] * N1[1][j] * G1[2][k];
          n++;
        }
  }
}

/* P3D_evaluate_geometry_elementQ2, PTAT3D_ELEMENT_OPERATION_STANDARD branch
 * (element_utils_q2.c:400-476; the STANDARD branch is the one compiled,
 * element_utils_q2.c:33). */
static void geom_q2(const double xe[81], double GNI[27][3][27], double detJ[27],
                    double dNx[27][27], double dNy[27][27], double dNz[27][27])
{
  for (int p = 0; p < 27; p++) {
    double J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    for (int k = 0; k < 27; k++)
      for (int a = 0; a < 3; a++)
        for (int d = 0; d < 3; d++) J[a][d] += GNI[p][a][k] * xe[3 * k + d];
    detJ[p] = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
              J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
              J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
    double t4 = J[2][0] * J[0][1], t6 = J[2][0] * J[0][2], t8 = J[1][0] * J[0][1];
    double t10 = J[1][0] * J[0][2], t12 = J[0][0] * J[1][1], t14 = J[0][0] * J[1][2];
    double t17 = 1.0 / (t4 * J[1][2] - t6 * J[1][1] - t8 * J[2][2] + t10 * J[2][1] +
                        t12 * J[2][2] - t14 * J[2][1]);
    double iJ[3][3];
    iJ[0][0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]) * t17;
    iJ[0][1] = -(J[0][1] * J[2][2] - J[0][2] * J[2][1]) * t17;
    iJ[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * t17;
    iJ[1][0] = -(-J[2][0] * J[1][2] + J[1][0] * J[2][2]) * t17;
    iJ[1][1] = (-t6 + J[0][0] * J[2][2]) * t17;
    iJ[1][2] = -(-t10 + t14) * t17;
    iJ[2][0] = (-J[2][0] * J[1][1] + J[1][0] * J[2][1]) * t17;
    iJ[2][1] = -(-t4 + J[0][0] * J[2][1]) * t17;
    iJ[2][2] = (-t8 + t12) * t17;
    for (int k = 0; k < 27; k++) {
      dNx[p][k] = iJ[0][0] * GNI[p][0][k] + iJ[0][1] * GNI[p][1][k] + iJ[0][2] * GNI[p][2][k];
      dNy[p][k] = iJ[1][0] * GNI[p][0][k] + iJ[1][1] * GNI[p][1][k] + iJ[1][2] * GNI[p][2][k];
      dNz[p][k] = iJ[2][0] * GNI[p][0][k] + iJ[2][1] * GNI[p][1][k] + iJ[2][2] * GNI[p][2][k];
    }
  }
}

static void gather_vec(const box_t *b, long e, const double *v, double out[81])
{
  long nid[27];
  elnodes(b, e, nid);
  for (int n = 0; n < 27; n++)
    for (int d = 0; d < 3; d++) out[3 * n + d] = v[3 * nid[n] + d];
}

static void scatter_add_vec(const box_t *b, long e, const double in[81], double *v)
{
  long nid[27];
  elnodes(b, e, nid);
  for (int n = 0; n < 27; n++)
    for (int d = 0; d < 3; d++) v[3 * nid[n] + d] += in[3 * n + d];
}

void orc_a11_ref(int mx, int my, int mz, const double *coords, const double *eta,
                 const double *u, double *y)
{
  box_t b = mkbox(mx, my, mz);
  static double W[27], XI[27][3], NI[27][27], GNI[27][3][27];
  prep_q2_3x3(W, XI, NI, GNI);
  for (long e = 0; e < b.nel; e++) {
    double xe[81], ue[81], Ye[81], detJ[27], dNx[27][27], dNy[27][27], dNz[27][27];
    gather_vec(&b, e, coords, xe);
    gather_vec(&b, e, u, ue);
    geom_q2(xe, GNI, detJ, dNx, dNy, dNz);
    memset(Ye, 0, sizeof Ye);
    for (int p = 0; p < 27; p++) {
      const double fac = W[p] * detJ[p], et = eta[27 * e + p];
      /* MatMultMF_Stokes_MixedFEM3d_B11 (stokes_q2p1_mf_operators_def.c:151-257):
       * engineering strain rate (exx,eyy,ezz,2exy,2exz,2eyz), tau = D eps */
      double sr[6] = {0, 0, 0, 0, 0, 0};
      for (int n = 0; n < 27; n++) {
        const double ux = ue[3 * n], uy = ue[3 * n + 1], uz = ue[3 * n + 2];
        sr[0] += ux * dNx[p][n];
        sr[1] += uy * dNy[p][n];
        sr[2] += uz * dNz[p][n];
        sr[3] += ux * dNy[p][n] + uy * dNx[p][n];
        sr[4] += ux * dNz[p][n] + uz * dNx[p][n];
        sr[5] += uy * dNz[p][n] + uz * dNy[p][n];
      }
      double t[6] = {2.0 * et * sr[0], 2.0 * et * sr[1], 2.0 * et * sr[2],
                     et * sr[3], et * sr[4], et * sr[5]};
      for (int n = 0; n < 27; n++) {
        Ye[3 * n + 0] += fac * (dNx[p][n] * t[0] + dNy[p][n] * t[3] + dNz[p][n] * t[4]);
        Ye[3 * n + 1] += fac * (dNx[p][n] * t[3] + dNy[p][n] * t[1] + dNz[p][n] * t[5]);
        Ye[3 * n + 2] += fac * (dNx[p][n] * t[4] + dNy[p][n] * t[5] + dNz[p][n] * t[2]);
      }
    }
    scatter_add_vec(&b, e, Ye, y);
  }
}

void orc_a11_element_matrix(int mx, int my, int mz, const double *coords, const double *eta,
                            long e, double *Ae)
{
  box_t b = mkbox(mx, my, mz);
  static double W[27], XI[27][3], NI[27][27], GNI[27][3][27];
  prep_q2_3x3(W, XI, NI, GNI);
  double xe[81], detJ[27], dNx[27][27], dNy[27][27], dNz[27][27];
  gather_vec(&b, e, coords, xe);
  geom_q2(xe, GNI, detJ, dNx, dNy, dNz);
  memset(Ae, 0, sizeof(double) * 81 * 81);
  for (int p = 0; p < 27; p++) {
    const double fac = W[p] * detJ[p], et = eta[27 * e + p];
    double Bm[6][81];
    memset(Bm, 0, sizeof Bm);
    for (int n = 0; n < 27; n++) {
      const double gx = dNx[p][n], gy = dNy[p][n], gz = dNz[p][n];
      Bm[0][3 * n] = gx;
      Bm[1][3 * n + 1] = gy;
      Bm[2][3 * n + 2] = gz;
      Bm[3][3 * n] = gy; Bm[3][3 * n + 1] = gx;
      Bm[4][3 * n] = gz; Bm[4][3 * n + 2] = gx;
      Bm[5][3 * n + 1] = gz; Bm[5][3 * n + 2] = gy;
    }
    const double dD[6] = {2 * fac * et, 2 * fac * et, 2 * fac * et, fac * et, fac * et, fac * et};
    for (int i = 0; i < 81; i++)
      for (int j = i; j < 81; j++) {
        double s = 0.0;
        for (int m = 0; m < 6; m++) s += Bm[m][i] * dD[m] * Bm[m][j];
        Ae[i * 81 + j] += s;
      }
  }
  for (int i = 0; i < 81; i++)
    for (int j = i + 1; j < 81; j++) Ae[j * 81 + i] = Ae[i * 81 + j];
}

void orc_a11_asm_matmult(int mx, int my, int mz, const double *coords, const double *eta,
                         const unsigned char *bcmask, const double *x, double *y)
{
  box_t b = mkbox(mx, my, mz);
  const long n3 = 3 * b.nnodes;
  memset(y, 0, sizeof(double) * n3);
  double *Ae = (double *)malloc(sizeof(double) * 81 * 81);
  for (long e = 0; e < b.nel; e++) {
    long nid[27];
    double xe[81], ye[81];
    elnodes(&b, e, nid);
    orc_a11_element_matrix(mx, my, mz, coords, eta, e, Ae);
    for (int n = 0; n < 27; n++)
      for (int d = 0; d < 3; d++) {
        long g = 3 * nid[n] + d;
        xe[3 * n + d] = bcmask[g] ? 0.0 : x[g]; /* masked column */
      }
    for (int i = 0; i < 81; i++) {
      double s = 0.0;
      for (int j = 0; j < 81; j++) s += Ae[i * 81 + j] * xe[j];
      ye[i] = s;
    }
    for (int n = 0; n < 27; n++)
      for (int d = 0; d < 3; d++) y[3 * nid[n] + d] += ye[3 * n + d];
  }
  free(Ae);
  for (long g = 0; g < n3; g++)
    if (bcmask[g]) y[g] = x[g]; /* BCListInsertScaling: unit diagonal */
}

void orc_a11_matmult(int mx, int my, int mz, const double *coords, const double *eta,
                     const unsigned char *bcmask, const double *x, double *y)
{
  box_t b = mkbox(mx, my, mz);
  const long n3 = 3 * b.nnodes;
  double *xl = (double *)malloc(sizeof(double) * n3);
  for (long g = 0; g < n3; g++) xl[g] = bcmask[g] ? 0.0 : x[g]; /* BCListInsertLocalZero */
  memset(y, 0, sizeof(double) * n3);
  orc_a11_tensor(mx, my, mz, coords, eta, xl, y);
  for (long g = 0; g < n3; g++)
    if (bcmask[g]) y[g] = x[g]; /* BCListInsertDirichlet_MatMult, scale 1 */
  free(xl);
}

void orc_a11_diagonal(int mx, int my, int mz, const double *coords, const double *eta,
                      const unsigned char *bcmask, double *dg)
{
  box_t b = mkbox(mx, my, mz);
  static double W[27], XI[27][3], NI[27][27], GNI[27][3][27];
  prep_q2_3x3(W, XI, NI, GNI);
  const long n3 = 3 * b.nnodes;
  memset(dg, 0, sizeof(double) * n3);
  for (long e = 0; e < b.nel; e++) {
    double xe[81], Ye[81], detJ[27], dNx[27][27], dNy[27][27], dNz[27][27];
    gather_vec(&b, e, coords, xe);
    geom_q2(xe, GNI, detJ, dNx, dNy, dNz);
    memset(Ye, 0, sizeof Ye);
    for (int p = 0; p < 27; p++) {
      const double fac = W[p] * detJ[p], et = eta[27 * e + p];
      /* MatMultMF_Stokes_MixedFEM3d_diagB11 (stokes_q2p1_mf_operators_diag_def.c:120-205) */
      for (int n = 0; n < 27; n++) {
        const double gx2 = dNx[p][n] * dNx[p][n], gy2 = dNy[p][n] * dNy[p][n],
                     gz2 = dNz[p][n] * dNz[p][n];
        Ye[3 * n + 0] += fac * (et * gy2 + et * gz2 + 2.0 * et * gx2);
        Ye[3 * n + 1] += fac * (et * gx2 + et * gz2 + 2.0 * et * gy2);
        Ye[3 * n + 2] += fac * (et * gx2 + et * gy2 + 2.0 * et * gz2);
      }
    }
    scatter_add_vec(&b, e, Ye, dg);
  }
  for (long g = 0; g < n3; g++)
    if (bcmask[g]) dg[g] = 1.0; /* BCListInsertValueIntoDirichletSlot(1.0) */
}

/* P3D_ConstructNi_P1GRel_3D (element_utils_q2.c:183-224): the P1disc basis
 * {1, (x-xbar)/Lx, (y-ybar)/Ly, (z-zbar)/Lz} at reference point xi. */
static void p1grel(const double xi[3], const double xe[81], double Np[4])
{
  double N1[3][3];
  for (int d = 0; d < 3; d++) {
    double t = xi[d];
    N1[d][0] = 0.5 * t * (t - 1.0);
    N1[d][1] = (1.0 + t) * (1.0 - t);
    N1[d][2] = 0.5 * (1.0 + t) * t;
  }
  double xg[3] = {0, 0, 0}, avg[3] = {0, 0, 0};
  int n = 0;
  for (int k = 0; k < 3; k++)
    for (int j = 0; j < 3; j++)
      for (int i = 0; i < 3; i++) {
        double Nn = N1[0][i] * N1[1][j] * N1[2][k];
        for (int d = 0; d < 3; d++) {
          xg[d] += Nn * xe[3 * n + d];
          avg[d] += xe[3 * n + d];
        }
        n++;
      }
  for (int d = 0; d < 3; d++) avg[d] = (1.0 / 27.0) * avg[d];
  const double L[3] = {xe[3 * 14] - xe[3 * 12], xe[3 * 16 + 1] - xe[3 * 10 + 1],
                       xe[3 * 22 + 2] - xe[3 * 4 + 2]};
  Np[0] = 1.0;
  for (int d = 0; d < 3; d++) Np[1 + d] = (xg[d] - avg[d]) / L[d];
}

void orc_a12_matmult(int mx, int my, int mz, const double *coords,
                     const unsigned char *bcmask, const double *p, double *yu)
{
  box_t b = mkbox(mx, my, mz);
  static double W[27], XI[27][3], NI[27][27], GNI[27][3][27];
  prep_q2_3x3(W, XI, NI, GNI);
  const long n3 = 3 * b.nnodes;
  memset(yu, 0, sizeof(double) * n3);
  for (long e = 0; e < b.nel; e++) {
    double xe[81], Ye[81], detJ[27], dNx[27][27], dNy[27][27], dNz[27][27];
    gather_vec(&b, e, coords, xe);
    geom_q2(xe, GNI, detJ, dNx, dNy, dNz);
    memset(Ye, 0, sizeof Ye);
    for (int q = 0; q < 27; q++) {
      double Np[4];
      p1grel(XI[q], xe, Np);
      const double fac = W[q] * detJ[q];
      /* MatMultMF_Stokes_MixedFEM3d_A12 (stokes_q2p1_mf_operators_def.c:423-516) */
      const double pq = Np[0] * p[4 * e] + Np[1] * p[4 * e + 1] + Np[2] * p[4 * e + 2] +
                        Np[3] * p[4 * e + 3];
      for (int n = 0; n < 27; n++) {
        Ye[3 * n + 0] += -fac * dNx[q][n] * pq;
        Ye[3 * n + 1] += -fac * dNy[q][n] * pq;
        Ye[3 * n + 2] += -fac * dNz[q][n] * pq;
      }
    }
    scatter_add_vec(&b, e, Ye, yu);
  }
  for (long g = 0; g < n3; g++)
    if (bcmask[g]) yu[g] = 0.0; /* BCListInsertZero (stokes_operators.c:1342) */
}

void orc_a21_matmult(int mx, int my, int mz, const double *coords,
                     const unsigned char *bcmask, const double *u, double *yp)
{
  box_t b = mkbox(mx, my, mz);
  static double W[27], XI[27][3], NI[27][27], GNI[27][3][27];
  prep_q2_3x3(W, XI, NI, GNI);
  memset(yp, 0, sizeof(double) * 4 * b.nel);
  for (long e = 0; e < b.nel; e++) {
    long nid[27];
    double xe[81], ue[81], detJ[27], dNx[27][27], dNy[27][27], dNz[27][27];
    elnodes(&b, e, nid);
    gather_vec(&b, e, coords, xe);
    for (int n = 0; n < 27; n++)
      for (int d = 0; d < 3; d++) {
        long g = 3 * nid[n] + d;
        ue[3 * n + d] = bcmask[g] ? 0.0 : u[g]; /* BCListInsertLocalZero (1450) */
      }
    geom_q2(xe, GNI, detJ, dNx, dNy, dNz);
    double Ye[4] = {0, 0, 0, 0};
    for (int q = 0; q < 27; q++) {
      double Np[4];
      p1grel(XI[q], xe, Np);
      const double fac = W[q] * detJ[q];
      /* MatMultMF_Stokes_MixedFEM3d_A21 (stokes_q2p1_mf_operators_def.c:517-538) */
      double dv = 0.0;
      for (int n = 0; n < 27; n++)
        dv += ue[3 * n] * dNx[q][n] + ue[3 * n + 1] * dNy[q][n] + ue[3 * n + 2] * dNz[q][n];
      for (int i = 0; i < 4; i++) Ye[i] += -fac * Np[i] * dv;
    }
    for (int i = 0; i < 4; i++) yp[4 * e + i] += Ye[i];
  }
}

void orc_schur_mass_blocks(int mx, int my, int mz, const double *coords, const double *eta,
                           double *Me)
{
  box_t b = mkbox(mx, my, mz);
  static double W[27], XI[27][3], NI[27][27], GNI[27][3][27];
  prep_q2_3x3(W, XI, NI, GNI);
  for (long e = 0; e < b.nel; e++) {
    double xe[81], detJ[27], dNx[27][27], dNy[27][27], dNz[27][27], A[16];
    gather_vec(&b, e, coords, xe);
    geom_q2(xe, GNI, detJ, dNx, dNy, dNz);
    memset(A, 0, sizeof A);
    for (int q = 0; q < 27; q++) {
      double Np[4];
      p1grel(XI[q], xe, Np);
      const double fac = (1.0 / eta[27 * e + q]) * W[q] * detJ[q];
      for (int i = 0; i < 4; i++)
        for (int j = i; j < 4; j++) A[i * 4 + j] -= fac * (Np[i] * Np[j]);
    }
    for (int i = 0; i < 4; i++)
      for (int j = i; j < 4; j++) A[j * 4 + i] = A[i * 4 + j];
    memcpy(Me + 16 * e, A, sizeof A);
  }
}

/* ------------------------------------------------------------------------ */
/* multigrid transfers                                                       */
/* ------------------------------------------------------------------------ */

/* 1-D lattice weights of the vertex-centred DMDA Q1 interpolation with
 * refinement 2: fine node I = 2i takes coarse node i with weight 1; fine node
 * 2i+1 takes coarse nodes i and i+1 with weight 1/2 each. */
static int lat_weights(long I, long *c, double *w)
{
  if ((I & 1) == 0) {
    c[0] = I / 2; w[0] = 1.0;
    return 1;
  }
  c[0] = I / 2; w[0] = 0.5;
  c[1] = I / 2 + 1; w[1] = 0.5;
  return 2;
}

void orc_prolong_add(int mxc, int myc, int mzc, const double *xc, double *yf)
{
  box_t bc = mkbox(mxc, myc, mzc), bf = mkbox(2 * mxc, 2 * myc, 2 * mzc);
  for (long K = 0; K < bf.Z; K++)
    for (long J = 0; J < bf.Y; J++)
      for (long I = 0; I < bf.X; I++) {
        long ci[2], cj[2], ck[2];
        double wi[2], wj[2], wk[2];
        int ni = lat_weights(I, ci, wi), nj = lat_weights(J, cj, wj), nk = lat_weights(K, ck, wk);
        long f = I + bf.X * (J + bf.Y * K);
        for (int c = 0; c < nk; c++)
          for (int bb = 0; bb < nj; bb++)
            for (int a = 0; a < ni; a++) {
              long g = ci[a] + bc.X * (cj[bb] + bc.Y * ck[c]);
              double w = wi[a] * wj[bb] * wk[c];
              for (int d = 0; d < 3; d++) yf[3 * f + d] += w * xc[3 * g + d];
            }
      }
}

void orc_restrict_add(int mxc, int myc, int mzc, const double *xf, double *yc)
{
  box_t bc = mkbox(mxc, myc, mzc), bf = mkbox(2 * mxc, 2 * myc, 2 * mzc);
  for (long K = 0; K < bf.Z; K++)
    for (long J = 0; J < bf.Y; J++)
      for (long I = 0; I < bf.X; I++) {
        long ci[2], cj[2], ck[2];
        double wi[2], wj[2], wk[2];
        int ni = lat_weights(I, ci, wi), nj = lat_weights(J, cj, wj), nk = lat_weights(K, ck, wk);
        long f = I + bf.X * (J + bf.Y * K);
        for (int c = 0; c < nk; c++)
          for (int bb = 0; bb < nj; bb++)
            for (int a = 0; a < ni; a++) {
              long g = ci[a] + bc.X * (cj[bb] + bc.Y * ck[c]);
              double w = wi[a] * wj[bb] * wk[c];
              for (int d = 0; d < 3; d++) yc[3 * g + d] += w * xf[3 * f + d];
            }
      }
}

void orc_inject_coords(int mxc, int myc, int mzc, const double *xf, double *xc)
{
  box_t bc = mkbox(mxc, myc, mzc), bf = mkbox(2 * mxc, 2 * myc, 2 * mzc);
  for (long k = 0; k < bc.Z; k++)
    for (long j = 0; j < bc.Y; j++)
      for (long i = 0; i < bc.X; i++) {
        long g = i + bc.X * (j + bc.Y * k), f = 2 * i + bf.X * (2 * j + bf.Y * 2 * k);
        for (int d = 0; d < 3; d++) xc[3 * g + d] = xf[3 * f + d];
      }
}

/* ------------------------------------------------------------------------ */
/* Chebyshev / Jacobi smoother (ksp_chebyrn.c:322-411)                       */
/* ------------------------------------------------------------------------ */
void orc_cheb_jacobi(int mx, int my, int mz, const double *coords, const double *eta,
                     const unsigned char *bcmask, const double *diag, double emin, double emax,
                     int its, int nonzero_guess, const double *b, double *x)
{
  box_t bx = mkbox(mx, my, mz);
  const long n = 3 * bx.nnodes;
  double *p[3], *r = (double *)malloc(sizeof(double) * n);
  p[0] = x;
  p[1] = (double *)malloc(sizeof(double) * n);
  p[2] = (double *)malloc(sizeof(double) * n);
  int km1 = 0, k = 1, kp1 = 2;
  const double scale = 2.0 / (emax + emin);
  const double alpha = 1.0 - scale * emin;
  const double mu = 1.0 / alpha, omegaprod = 2.0 / alpha;
  double c[3];
  c[km1] = 1.0;
  c[k] = mu;
  if (nonzero_guess) {
    orc_a11_matmult(mx, my, mz, coords, eta, bcmask, x, r);
    for (long i = 0; i < n; i++) r[i] = b[i] - r[i];
  } else {
    memset(x, 0, sizeof(double) * n);
    memcpy(r, b, sizeof(double) * n);
  }
  /* PCJACOBI applies the stored reciprocal of the diagonal */
  for (long i = 0; i < n; i++) p[k][i] = x[i] + scale * (r[i] * (1.0 / diag[i]));
  for (int it = 0; it < its; it++) {
    c[kp1] = 2.0 * mu * c[k] - c[km1];
    const double omega = omegaprod * c[k] / c[kp1];
    orc_a11_matmult(mx, my, mz, coords, eta, bcmask, p[k], r);
    for (long i = 0; i < n; i++) r[i] = b[i] - r[i];
    for (long i = 0; i < n; i++) {
      double z = r[i] * (1.0 / diag[i]);
      p[kp1][i] = omega * scale * z + (1.0 - omega) * p[km1][i] + omega * p[k][i];
    }
    int t = km1;
    km1 = k; k = kp1; kp1 = t;
  }
  if (k != 0) memcpy(x, p[k], sizeof(double) * n);
  free(r);
  free(p[1]);
  free(p[2]);
}

/* ------------------------------------------------------------------------ */
/* strain-rate-dependent viscosity at the quadrature points                  */
/* ------------------------------------------------------------------------ */

/* eta[27e+q] = clip(B[27e+q] * max(eII, 1e-17)^(1/n - 1), eta_min, eta_max),
 * eII = sqrt(0.5 D:D), D = sym(grad u) at the quadrature point
 * (ComputeStrainRate3d / ComputeSecondInvariant3d, stokes_rheology_vp_std.c:85-148;
 * the power law of VISCOUS_ARRHENIUS_2, stokes_rheology_vp_std.c:437-461, with
 * the prefactor B = Ascale preexpA^(-1/n) exp(E / (n R T)) given per point). */
void orc_powerlaw_eta(int mx, int my, int mz, const double *coords, const double *u,
                      const double *B, double nexp, double eta_min, double eta_max, double *eta)
{
  box_t b = mkbox(mx, my, mz);
  static double W[27], XI[27][3], NI[27][27], GNI[27][3][27];
  prep_q2_3x3(W, XI, NI, GNI);
  for (long e = 0; e < b.nel; e++) {
    double xe[81], ue[81], detJ[27], dNx[27][27], dNy[27][27], dNz[27][27];
    gather_vec(&b, e, coords, xe);
    gather_vec(&b, e, u, ue);
    geom_q2(xe, GNI, detJ, dNx, dNy, dNz);
    for (int q = 0; q < 27; q++) {
      double exx = 0, eyy = 0, ezz = 0, exy = 0, exz = 0, eyz = 0;
      for (int k = 0; k < 27; k++) {
        const double ux = ue[3 * k], uy = ue[3 * k + 1], uz = ue[3 * k + 2];
        exx += dNx[q][k] * ux;
        eyy += dNy[q][k] * uy;
        ezz += dNz[q][k] * uz;
        exy += dNy[q][k] * ux + dNx[q][k] * uy;
        exz += dNz[q][k] * ux + dNx[q][k] * uz;
        eyz += dNz[q][k] * uy + dNy[q][k] * uz;
      }
      exy *= 0.5;
      exz *= 0.5;
      eyz *= 0.5;
      const double sum = exx * exx + eyy * eyy + ezz * ezz + 2.0 * (exy * exy + exz * exz + eyz * eyz);
      double sr = sqrt(0.5 * sum);
      if (sr < 1.0e-17) sr = 1.0e-17;
      double v = B[27 * e + q] * pow(sr, 1.0 / nexp - 1.0);
      if (v < eta_min) v = eta_min;
      if (v > eta_max) v = eta_max;
      eta[27 * e + q] = v;
    }
  }
}


/* ------------------------------------------------------------------------ */
/* coefficient hierarchy: Q1 projection of quadrature-point viscosity        */
/* ------------------------------------------------------------------------ */

/* weight of the qp at reference point xi for local node n of a Q2 element in
 * the projection: the Q1 basis of the corner for the 8 corner nodes, 1 for
 * the 19 others (material_point_utils.c:1190-1210) */
static void q1proj_weights(const double xi[3], double w[27])
{
  for (int n = 0; n < 27; n++) w[n] = 1.0;
  for (int c = 0; c < 2; c++)
    for (int b = 0; b < 2; b++)
      for (int a = 0; a < 2; a++) {
        const double fa = a ? 0.5 * (1.0 + xi[0]) : 0.5 * (1.0 - xi[0]);
        const double fb = b ? 0.5 * (1.0 + xi[1]) : 0.5 * (1.0 - xi[1]);
        const double fc = c ? 0.5 * (1.0 + xi[2]) : 0.5 * (1.0 - xi[2]);
        w[2 * a + 6 * b + 18 * c] = fa * fb * fc;
      }
}

/* Fine-level nodal field from the quadrature-point values (element-major
 * eta[27e+q]): nodal[n] = sum_{e,q} w_q(n) eta_{e,q} / sum_{e,q} w_q(n)
 * (_BuildQ1CoefficientProjection_QuadraturePoints_MPntPStokes_FineGrid,
 * material_point_utils.c:1138-1265; element loop, qp loop, then the
 * pointwise division). */
void orc_q1_nodal_from_qp(int mx, int my, int mz, const double *eta, double *nodal)
{
  box_t b = mkbox(mx, my, mz);
  static double W[27], XI[27][3], NI[27][27], GNI[27][3][27];
  prep_q2_3x3(W, XI, NI, GNI);
  double *B = (double *)calloc((size_t)b.nnodes, sizeof(double));
  memset(nodal, 0, sizeof(double) * b.nnodes);
  for (long e = 0; e < b.nel; e++) {
    double Ae[27] = {0}, Be[27] = {0};
    long nid[27];
    elnodes(&b, e, nid);
    for (int q = 0; q < 27; q++) {
      double w[27];
      q1proj_weights(XI[q], w);
      for (int n = 0; n < 27; n++) {
        Ae[n] += w[n] * eta[27 * e + q];
        Be[n] += w[n];
      }
    }
    for (int n = 0; n < 27; n++) {
      nodal[nid[n]] += Ae[n];
      B[nid[n]] += Be[n];
    }
  }
  for (long n = 0; n < b.nnodes; n++) nodal[n] /= B[n];
  free(B);
}

/* One level down the hierarchy on the node lattice (scalar DMDA Q1
 * interpolation P between the coarse box and the fine box 2mxc x 2myc x 2mzc):
 * inject == 0: nc = (P^T nf) * (1 / (P^T 1)) (MatRestrict and
 * DMCreateInterpolationScale, material_point_utils.c:1373-1386);
 * inject != 0: nc = nf at the even fine nodes (DMCreateInjection, 1389-1395). */
void orc_q1_restrict_scaled(int mxc, int myc, int mzc, int inject, const double *nf, double *nc)
{
  box_t bc = mkbox(mxc, myc, mzc), bf = mkbox(2 * mxc, 2 * myc, 2 * mzc);
  if (inject) {
    for (long k = 0; k < bc.Z; k++)
      for (long j = 0; j < bc.Y; j++)
        for (long i = 0; i < bc.X; i++) nc[i + bc.X * (j + bc.Y * k)] = nf[2 * i + bf.X * (2 * j + bf.Y * 2 * k)];
    return;
  }
  double *s = (double *)calloc((size_t)bc.nnodes, sizeof(double));
  memset(nc, 0, sizeof(double) * bc.nnodes);
  for (long K = 0; K < bf.Z; K++)
    for (long J = 0; J < bf.Y; J++)
      for (long I = 0; I < bf.X; I++) {
        long ci[2], cj[2], ck[2];
        double wi[2], wj[2], wk[2];
        int ni = lat_weights(I, ci, wi), nj = lat_weights(J, cj, wj), nk = lat_weights(K, ck, wk);
        const long f = I + bf.X * (J + bf.Y * K);
        for (int c = 0; c < nk; c++)
          for (int bb = 0; bb < nj; bb++)
            for (int a = 0; a < ni; a++) {
              const long g = ci[a] + bc.X * (cj[bb] + bc.Y * ck[c]);
              const double w = wi[a] * wj[bb] * wk[c];
              nc[g] += w * nf[f];
              s[g] += w;
            }
      }
  for (long g = 0; g < bc.nnodes; g++) nc[g] *= 1.0 / s[g];
  free(s);
}

/* Quadrature-point values of a level from its nodal field: Q1 interpolation
 * from the 8 element corners, clamped to the range of the element's 27 nodal
 * values (_SwarmUpdateGaussPropertiesLocalL2ProjectionQ1_MPntPStokes_Interpolate
 * ToQuadratePoints, material_point_utils.c:894-1000). */
void orc_q1_to_qp(int mx, int my, int mz, const double *nodal, double *eta)
{
  box_t b = mkbox(mx, my, mz);
  static double W[27], XI[27][3], NI[27][27], GNI[27][3][27];
  prep_q2_3x3(W, XI, NI, GNI);
  for (long e = 0; e < b.nel; e++) {
    long nid[27];
    elnodes(&b, e, nid);
    double lo = 1.0e32, hi = -1.0e32;
    for (int n = 0; n < 27; n++) {
      const double v = nodal[nid[n]];
      if (v < lo) lo = v;
      if (v > hi) hi = v;
    }
    for (int q = 0; q < 27; q++) {
      double w[27], v = 0.0;
      q1proj_weights(XI[q], w);
      for (int c = 0; c < 2; c++)
        for (int bb = 0; bb < 2; bb++)
          for (int a = 0; a < 2; a++) {
            const int n = 2 * a + 6 * bb + 18 * c;
            v += w[n] * nodal[nid[n]];
          }
      if (v < lo) v = lo;
      if (v > hi) v = hi;
      eta[27 * e + q] = v;
    }
  }
}

/* Marker version of the fine nodal field: point p of element elem[p] at
 * reference coordinates xi[3p..3p+2] carries val[p]; same weights as the
 * quadrature-point version, accumulated point by point
 * (_SwarmUpdateGaussPropertiesLocalL2ProjectionQ1_MPntPStokes,
 * material_point_utils.c:604-705). */
void orc_q1_nodal_from_points(int mx, int my, int mz, long npts, const long *elem, const double *xi,
                              const double *val, double *nodal)
{
  box_t b = mkbox(mx, my, mz);
  double *B = (double *)calloc((size_t)b.nnodes, sizeof(double));
  memset(nodal, 0, sizeof(double) * b.nnodes);
  for (long p = 0; p < npts; p++) {
    double w[27];
    long nid[27];
    q1proj_weights(&xi[3 * p], w);
    elnodes(&b, elem[p], nid);
    for (int n = 0; n < 27; n++) {
      nodal[nid[n]] += w[n] * val[p];
      B[nid[n]] += w[n];
    }
  }
  for (long n = 0; n < b.nnodes; n++) nodal[n] /= B[n];
  free(B);
}
