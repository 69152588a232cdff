/*
 * ptmf_oracle.h -- CPU restatement of pTatin3d's matrix-free Q2-P1disc Stokes
 * operator stack, used ONLY as a correctness checker and CPU baseline.
 *
 * TEST INFRASTRUCTURE: nothing in the product library links or calls this.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.
 *
 * Mesh model: a single structured box of mx*my*mz Q2 elements with
 * X=2mx+1, Y=2my+1, Z=2mz+1 nodes.  Node (i,j,k) has index i + X*(j + Y*k)
 * (dmda_element_q2p1.c:395-440); vectors are node-interlaced, dof 3*node+d.
 * Element e = ei + mx*(ej + my*ek); local node n = i + 3j + 9k.
 * Quadrature-point data (viscosity) is element-major: eta[27*e + q] with
 * q = qi + 3qj + 9qk (qi <-> xi).  Pressure is P1disc: p[4*e + i].
 */
#ifndef PTMF_ORACLE_H
#define PTMF_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

/* number of OpenMP threads the oracle runs on (1 if built without OpenMP) */
int orc_num_threads(void);
void orc_set_num_threads(int n);

/* y += A11_local * u, no boundary conditions.  Sum-factorized tensor path,
 * NEV=4 element batching, OpenMP element loop with a serialized scatter:
 * restates MFStokesWrapper_A11_Tensor (stokes_operators_tensor.c:181-288). */
void orc_a11_tensor(int mx, int my, int mz, const double *coords,
                    const double *eta, const double *u, double *y);

/* y += A11_local * u via per-quadrature-point dense element kernels with the
 * truncated Gauss constants: restates MFStokesWrapper_A11 "ref"
 * (stokes_operators_mf.c:233-347) + MatMultMF_Stokes_MixedFEM3d_B11
 * (stokes_q2p1_mf_operators_def.c:151-257). */
void orc_a11_ref(int mx, int my, int mz, const double *coords,
                 const double *eta, const double *u, double *y);

/* 81x81 element stiffness B^T diag(2eta,2eta,2eta,eta,eta,eta) w detJ B of
 * element e: restates stokes_assembly.c:219-268.  Row/col index 3*n+d. */
void orc_a11_element_matrix(int mx, int my, int mz, const double *coords,
                            const double *eta, long e, double *Ae);

/* y = A_asm x with Dirichlet rows/cols removed and unit diagonal: the
 * assembled operator of MatAssemble_StokesA_AUU (stokes_assembly.c:115-297)
 * applied element by element. */
void orc_a11_asm_matmult(int mx, int my, int mz, const double *coords,
                         const double *eta, const unsigned char *bcmask,
                         const double *x, double *y);

/* MatMult_MFStokes_A11 (stokes_operators.c:1142-1205) on top of the tensor
 * kernel: y = A11 * mask(x); y[dir] = x[dir]. */
void orc_a11_matmult(int mx, int my, int mz, const double *coords,
                     const double *eta, const unsigned char *bcmask,
                     const double *x, double *y);

/* MatGetDiagonal_MFStokes_A11 (stokes_operators.c:1060-1099) via
 * MFStokesWrapper_diagA11 (stokes_operators_mf.c:67-131). */
void orc_a11_diagonal(int mx, int my, int mz, const double *coords,
                      const double *eta, const unsigned char *bcmask,
                      double *d);

/* MatMult_MFStokes_A12 (stokes_operators.c:1284-1350): yu = B p, Dirichlet
 * velocity rows zeroed.  p has 4 dofs per element. */
void orc_a12_matmult(int mx, int my, int mz, const double *coords,
                     const unsigned char *bcmask, const double *p, double *yu);

/* MatMult_MFStokes_A21 (stokes_operators.c:1424-1486): yp = B^T mask(u). */
void orc_a21_matmult(int mx, int my, int mz, const double *coords,
                     const unsigned char *bcmask, const double *u, double *yp);

/* Scaled viscous mass matrix Schur approximation diagonal blocks:
 * restates MatAssemble_StokesPC_ScaledMassMatrix (stokes_assembly.c:299-450):
 * M_e[i][j] = -sum_q w detJ / eta_q NIp_i NIp_j, 16 values per element. */
void orc_schur_mass_blocks(int mx, int my, int mz, const double *coords,
                           const double *eta, double *Me);

/* Geometric-multigrid transfer between a coarse box (mxc,myc,mzc) and the
 * fine box (2mxc,2myc,2mzc): Q1 interpolation on the node lattice with block
 * size 3 (PETSc DMCreateInterpolation for DMDA, called at
 * ptatin_driver_linear_ts.c:989).  prolong: yf += P xc; restrict: yc += P^T xf */
void orc_prolong_add(int mxc, int myc, int mzc, const double *xc, double *yf);
void orc_restrict_add(int mxc, int myc, int mzc, const double *xf, double *yc);

/* Coordinate injection fine -> coarse (dmda_project_coords.c:39-71). */
void orc_inject_coords(int mxc, int myc, int mzc, const double *xf, double *xc);

/* Chebyshev iteration with Jacobi preconditioning on the BC-aware A11
 * operator: restates the recurrence of KSPSolve_ChebychevRN_lag_norm
 * (ksp_chebyrn.c:322-411) with no norm checks.  x holds the initial guess
 * (nonzero_guess != 0) and is overwritten with the result. */
void orc_cheb_jacobi(int mx, int my, int mz, const double *coords,
                     const double *eta, const unsigned char *bcmask,
                     const double *diag, double emin, double emax, int its,
                     int nonzero_guess, const double *b, double *x);

/* eta[27e+q] = clip(B[27e+q] * max(eII, 1e-17)^(1/n-1), eta_min, eta_max),
 * eII = sqrt(0.5 D:D) of u at the quadrature points (stokes_rheology_vp_std.c:85-148,437-461) */
void orc_powerlaw_eta(int mx, int my, int mz, const double *coords, const double *u,
                      const double *B, double nexp, double eta_min, double eta_max, double *eta);

/* Coefficient hierarchy from quadrature-point viscosity
 * (SwarmUpdateGaussPropertiesLocalL2Projection_Q1_MPntPStokes_Hierarchy,
 * material_point_utils.c:1267-1420, quadrature-point projection types):
 *   orc_q1_nodal_from_qp    fine nodal field (corner nodes Q1-weighted, other
 *                           nodes weight 1), nodal[(2mx+1)(2my+1)(2mz+1)];
 *   orc_q1_restrict_scaled  next coarser nodal field, (P^T f) / (P^T 1) or
 *                           injection (inject != 0);
 *   orc_q1_to_qp            Q1 interpolation to the 27 qps, clamped to the
 *                           element's nodal range, element-major eta[27e+q]. */
void orc_q1_nodal_from_qp(int mx, int my, int mz, const double *eta, double *nodal);
void orc_q1_restrict_scaled(int mxc, int myc, int mzc, int inject, const double *nf, double *nc);
void orc_q1_to_qp(int mx, int my, int mz, const double *nodal, double *eta);
/* The fine nodal field from material points instead of quadrature points:
 * point p in element elem[p] at reference coordinates xi[3p..3p+2] with value
 * val[p] (_SwarmUpdateGaussPropertiesLocalL2ProjectionQ1_MPntPStokes,
 * material_point_utils.c:604-705). */
void orc_q1_nodal_from_points(int mx, int my, int mz, long npts, const long *elem, const double *xi,
                              const double *val, double *nodal);

#ifdef __cplusplus
}
#endif
#endif
