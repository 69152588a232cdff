/*
 * ptmf.h -- C-ABI of the MI355X-native matrix-free Q2-P1disc Stokes operator
 * stack (gfx950).  Plain C types only: pointers, sizes, status codes.
 *
 * A PETSc host keeps its own objects; these entry points replace the bodies
 * of the reference's operator plugin and MatShell operations:
 *
 *   reference interface (pTatin3d)                          ptmf entry point
 *   ------------------------------------------------------  -----------------------------
 *   MatA11MFCreate/MatA11MFSetup, SpMVOp_SetUp               ptmf_op_create + ptmf_op_set_*
 *     (stokes_operators.c:101-157,226-273; stokes_operators.h:47-62)
 *   SpMVOp_Destroy / MFA11Destroy_CUDA (cuda.cu:290-323)     ptmf_op_destroy
 *   SpMVOp_MatMult = MFStokesWrapper_A11_{Tensor,AVX,CUDA}   ptmf_a11_spmv[_host]
 *     (stokes_operators_tensor.c:181; stokes_operators_mf.h:36-66)
 *   MatMult_MFStokes_A11 (stokes_operators.c:1142-1205)      ptmf_a11_mult[_f32]
 *   MatGetDiagonal_MFStokes_A11 (stokes_operators.c:1060)    ptmf_a11_diagonal
 *   MatMult_MFStokes_A12 (stokes_operators.c:1284-1350)      ptmf_a12_mult
 *   MatMult_MFStokes_A21 (stokes_operators.c:1424-1486)      ptmf_a21_mult
 *
 * Array conventions (identical to the reference's local vectors on one rank):
 *   coords, velocity: 3*N doubles, node-interlaced, node (i,j,k) of the Q2
 *     lattice -> i + X*(j + Y*k), X=2mx+1, Y=2my+1 (dmda_element_q2p1.c:395-440)
 *   eta: 27*E doubles, element-major, qp q = qi + 3qj + 9qk
 *     (QPntVolCoefStokes.eta of VolumeQuadratureGetCellData_Stokes,
 *      ptatin3d_stokes.c:486-508; element e = ei + mx*(ej + my*ek))
 *   pressure: 4*E doubles, p[4e+i], P1disc basis {1,x,y,z} relative to the
 *     element centre (element_utils_q2.c:183-224)
 *   Dirichlet mask: 3*N bytes, nonzero = Dirichlet dof (BCList_DIRICHLET).
 *
 * Unless the name ends in _host, vector arguments are DEVICE pointers
 * (allocate with ptmf_malloc).  All work is ordered on the operator's
 * stream; entry points that return host-visible results synchronize it, and
 * the ptmf_memcpy_* / ptmf_memset helpers wait for the whole device first.
 * Every function returns PTMF_OK (0) or an error code; ptmf_last_error()
 * gives the message.
 */
#ifndef PTMF_H
#define PTMF_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PTMF_OK 0
#define PTMF_ERR_ARG 1        /* bad argument (sizes, null pointers)            */
#define PTMF_ERR_NOT_SETUP 2  /* "MF operator not setup" (stokes_operators.c:1160) */
#define PTMF_ERR_DEVICE 3     /* HIP runtime error                              */
#define PTMF_ERR_NOMEM 4      /* device allocation failed                       */
#define PTMF_ERR_UNSUPPORTED 5 /* unknown option value (cf. "No -a11_op %s", :151) */
#define PTMF_ERR_DIVERGED 6   /* Krylov/MG iteration produced NaN/Inf            */

typedef struct ptmf_op_s *ptmf_op;
typedef struct ptmf_mg_s *ptmf_mg;
typedef struct ptmf_batch_s *ptmf_batch;
typedef struct ptmf_bmg_s *ptmf_bmg;

/* Options of the geometric-multigrid preconditioner of the viscous block
 * (the reference configures the same solver through PETSc options:
 * -fieldsplit_u_mg_levels_ksp_type chebyshev, ..._esteig, ..._max_it,
 * ..._pc_type jacobi, -fieldsplit_u_mg_coarse_pc_type lu; see
 * pTatin3dStokesKSPConfigureFSGMG, ptatin_driver_linear_ts.c:819-882). */
typedef struct {
  int levels_mg;          /* finest levels in the V-cycle (0: all); coarser levels form
                             the coarse solver (FP64 PCG + FP64 V-cycle, exact at level 0) */
  int smoother_precision; /* 32 or 64: operator, Chebyshev and transfers in the V-cycle */
  int pre_its, post_its;  /* Chebyshev iterations per smoothing (KSP max_it) */
  double esteig[4];       /* emin = e0*min + e1*max, emax = e2*min + e3*max of D^-1 A
                             (KSPChebyshevEstEigSet; default 0, 0.2, 0, 1.1) */
  int esteig_its;         /* Lanczos (Jacobi-PCG) steps of the eigenvalue estimate */
  double coarse_rtol;     /* relative tolerance of the coarse PCG */
  int coarse_maxit;
  int galerkin_levels;    /* coarsest levels whose operator is the Galerkin product
                             R A_{l+1} P (OP_TYPE_GALERKIN of -A11_operator_type,
                             ptatin_driver_linear_ts.c:712-793) instead of the
                             rediscretised matrix-free operator; 0: none */
  int bc_transfer;        /* 1: the coarse correction vanishes on Dirichlet dofs (the
                             restricted residual is zeroed there); 0: plain R = P^T as
                             PCMG with DMDA interpolation (ptatin_driver_linear_ts.c:849) */
  int smoother_diag;      /* 0: point Jacobi (diag(A), -mg_levels_pc_type jacobi); 1: l1-Jacobi
                             (sum_j |a_ij|, which bounds the spectrum of D^-1 A by 1 where
                             point Jacobi lets elements cut by viscosity jumps stretch it) */
  int use_graph;          /* 1: replay the V-cycle as a captured HIP graph from its second call
                             (needs levels_mg == nlevels; a new ptmf_mg_setup or
                             ptmf_mg_set_level_eigs recaptures, and so does a change of a level
                             operator's coefficients, tiling, kernel variant or Dirichlet list
                             through ptmf_mg_level_op: the next cycle runs eagerly, the one
                             after it is captured again) */
  int transfer;           /* 0: DMDA Q1 interpolation on the Q2 node lattice (the reference's
                             DMCreateInterpolation, ptatin_driver_linear_ts.c:989); 1: Q2
                             embedding (fine nodal values of the coarse Q2 function) */
  int coarse_factor;      /* exact solve of level 0 (the reference assembles it,
                             MatAssemble_StokesA_AUU, stokes_assembly.c:115-297, and factors it,
                             -fieldsplit_u_mg_coarse_pc_type lu): 1 dense explicit inverse (at most
                             32,768 dofs; needed by the batched multigrid); 2 sparse multifrontal
                             Cholesky of the assembled level (nested-dissection order, any size);
                             0 (default) dense up to 32,768 dofs, sparse above */
} ptmf_mg_options;

/* ---- library / device helpers ------------------------------------------ */
const char *ptmf_last_error(void);
int ptmf_version(int *major, int *minor);
int ptmf_device_count(int *n);
int ptmf_set_device(int dev);
int ptmf_malloc(void **dptr, size_t bytes);
int ptmf_free(void *dptr);
int ptmf_memcpy_htod(void *dst, const void *src, size_t bytes);
int ptmf_memcpy_dtoh(void *dst, const void *src, size_t bytes);
/* device-to-device copy with a plain streaming kernel (16-byte aligned) */
int ptmf_memcpy_dtod(void *dst, const void *src, size_t bytes);
int ptmf_memset(void *dptr, int value, size_t bytes);
int ptmf_device_synchronize(void);

/* ---- operator lifecycle (MatA11MFCreate / MatA11MFSetup) ----------------- */
/* Creates the velocity-block operator of an mx*my*mz Q2 box on the current
 * device.  Coordinates, viscosity and the Dirichlet mask must be set before
 * any apply (else PTMF_ERR_NOT_SETUP). */
int ptmf_op_create(int mx, int my, int mz, ptmf_op *op);
int ptmf_op_destroy(ptmf_op *op);
int ptmf_op_sizes(ptmf_op op, long *nnodes, long *nel);
/* host arrays, copied to the device (coords 3N, eta 27E, mask 3N) */
int ptmf_op_set_coordinates(ptmf_op op, const double *coords);
int ptmf_op_set_viscosity(ptmf_op op, const double *eta);
int ptmf_op_set_dirichlet(ptmf_op op, const unsigned char *mask);
/* device-array variants (no host round trip) */
int ptmf_op_set_coordinates_device(ptmf_op op, const double *coords);
int ptmf_op_set_viscosity_device(ptmf_op op, const double *eta);
/* Strain-rate-dependent viscosity at the quadrature points (device arrays,
 * element-major 27E like ptmf_op_set_viscosity):
 *   eta = clip(B * max(eII, 1e-17)^(1/n - 1), eta_min, eta_max),
 *   eII = sqrt(0.5 D:D), D = sym(grad u)
 * (ComputeStrainRate3d / ComputeSecondInvariant3d and the VISCOUS_ARRHENIUS_2
 * law, stokes_rheology_vp_std.c:85-148,437-461, evaluated at quadrature points
 * with B = Ascale preexpA^(-1/n) exp(E/(n R T)) given per point).  Feed the
 * result to ptmf_op_set_viscosity_device for a Picard update. */
int ptmf_op_powerlaw_viscosity(ptmf_op op, const double *u, const double *B, double nexp, double eta_min,
                               double eta_max, double *eta);
/* Material-point viscosity at the quadrature points of op's box: point p
 * lies in element elem[p] (64-bit) at reference coordinates xi[3p..3p+2] with
 * value val[p].  Nodal field = sum_p w_p val_p / sum_p w_p with the Q1 corner
 * basis at the 8 corners and weight 1 at the other 19 nodes, then Q1
 * interpolation to the quadrature points clamped to the element's nodal range
 * (SwarmUpdateGaussPropertiesLocalL2Projection_Q1_MPntPStokes,
 * material_point_utils.c:604-705,842-890).  Device arrays; eta is 27E
 * element-major, ready for ptmf_op_set_viscosity_device. */
int ptmf_op_project_points(ptmf_op op, long npts, const long *elem, const double *xi, const double *val,
                           double *eta);
/* Work decomposition of the MatMult kernel: TX x TY element tiles (TX*TY <= 21)
 * marching through ZC element layers.  A default is chosen at creation; the
 * autotuner overrides it.  No reference counterpart (the CUDA backend's
 * coloring, stokes_operators_cuda.cu:337-397, plays this role there). */
int ptmf_op_set_tiling(ptmf_op op, int TX, int TY, int ZC);
int ptmf_op_get_tiling(ptmf_op op, int *TX, int *TY, int *ZC);
/* MatMult kernel variant: 0 = z-march; 1 = z-march with the next layer's
 * loads issued one stage ahead (software pipelined; equal to 0 to rounding);
 * 2 = stored quadrature-point geometry (partial assembly: 9 values per
 * quadrature point, rebuilt after a coefficient or tiling change; same
 * operator to rounding, needs positive 2 w eta / det); 3 = coloured scatter
 * (one launch per element parity class, sums added into y directly, no tiles
 * or face buffers; the colouring of the reference's CUDA backend,
 * stokes_operators_cuda.cu:337-397, with the 3-lane element kernel) */
int ptmf_op_set_kernel_variant(ptmf_op op, int variant);
int ptmf_op_get_kernel_variant(ptmf_op op, int *variant);
/* Elements per work-group of the variant-0 MatMult: 1, 2 or 4 tiles (one
 * 64-lane wave and TX*TY elements each) per work-group; the autotuner drives
 * it.  No reference counterpart (launch shape of the backend). */
int ptmf_op_set_tiles_per_group(ptmf_op op, int tiles);
int ptmf_op_get_tiles_per_group(ptmf_op op, int *tiles);
/* coefficient state counter, bumped by every set_* call (replaces the
 * PetscObjectState stamp of stokes_operators.c:1162-1163) */
int ptmf_op_state(ptmf_op op, unsigned long *state);

/* ---- A11 (viscous block) ------------------------------------------------- */
/* y += A11_local u with no boundary handling: the SpMVOp_MatMult contract
 * (stokes_operators.h:59; MFStokesWrapper_A11_Tensor, stokes_operators_tensor.c:181). */
int ptmf_a11_spmv(ptmf_op op, const double *u, double *y);
/* Same contract on HOST arrays (for MFStokesWrapper_A11_HIP in a PETSc
 * adapter): copies u in and writes Yu = A11_local u.  The reference's caller
 * zeroes Yu right before the plugin call (stokes_operators.c:1181), so this
 * is its Yu += A u; like the CUDA backend (stokes_operators_cuda.cu:453) the
 * old contents of Yu are not read, which saves one PCIe transfer. */
int ptmf_a11_spmv_host(ptmf_op op, const double *u, double *Yu);
/* MatMult_MFStokes_A11: y = A11 * mask(x) on free rows, y = x on Dirichlet rows.
 * Asynchronous: returns once the work is enqueued on the operator's stream;
 * later work of the same operator, the ptmf_memcpy_* / ptmf_memset helpers
 * and ptmf_op_synchronize are ordered after it. */
int ptmf_a11_mult(ptmf_op op, const double *x, double *y);
/* FP32 version of ptmf_a11_mult (FP32 coordinates, viscosity and accumulation). */
int ptmf_a11_mult_f32(ptmf_op op, const float *x, float *y);
/* ptmf_a11_mult that allocates nothing (no hipMalloc / hipFree / allocating
 * library call).  No counterpart in the reference: it is for callers that
 * cannot allocate (code under HIP graph capture, tight serving loops,
 * harnesses that forbid allocation).  The z-march's face buffers live in a
 * static device array of 64 MiB per GPU (8,388,608 doubles, held until the
 * process ends): a box and tiling whose face buffers need more return
 * PTMF_ERR_UNSUPPORTED.  Same kernels and results as ptmf_a11_mult (the
 * stored-geometry and coloured kernel variants run as the default z-march).
 * The operator must be set up (ptmf_op_set_viscosity / ptmf_op_set_tiling
 * allocate its tile-ordered viscosity) and belong to the current device.
 * All calls on one GPU share the array: calls that overlap on one GPU race. */
int ptmf_a11_mult_noalloc(ptmf_op op, const double *x, double *y);
/* wait for all work enqueued on the operator's stream */
int ptmf_op_synchronize(ptmf_op op);
/* MatGetDiagonal_MFStokes_A11: diag(A11), 1 in Dirichlet slots. */
int ptmf_a11_diagonal(ptmf_op op, double *d);
/* l1 row sums of the A11 MatMult operator (device vector 3N): for each
 * non-Dirichlet row, the sum over elements of the absolute entries of the
 * element matrix in non-Dirichlet columns (an upper bound of sum_j |a_ij|);
 * Dirichlet rows 1.  The scaling of l1-Jacobi smoothing; no counterpart in
 * the reference, whose smoother is point Jacobi (PCJACOBI). */
int ptmf_a11_l1_diagonal(ptmf_op op, double *d);

/* ---- B / B^T ------------------------------------------------------------- */
/* MatMult_MFStokes_A12: yu = B p (p: 4E), Dirichlet velocity rows set to 0. */
int ptmf_a12_mult(ptmf_op op, const double *p, double *yu);
/* MatMult_MFStokes_A21: yp = B^T mask(u) (yp: 4E). */
int ptmf_a21_mult(ptmf_op op, const double *u, double *yp);
/* MatMult_MFStokes_A (stokes_operators.c:884-967, the fused momentum +
 * continuity wrapper MFStokesWrapper_A): x = [u (3N) | p (4E)], y = K x with
 * K = [A11 A12; A21 0], Dirichlet velocity columns ignored and Dirichlet
 * velocity rows y_u = x_u.  One pass over the elements (the pressure enters
 * the A11 stress at each quadrature point). */
int ptmf_stokes_mult(ptmf_op op, const double *x, double *y);
/* Me[16e + 4i + j] = -sum_q w detJ / eta_q NIp_i NIp_j: the element blocks of
 * the scaled viscous pressure mass matrix used as Schur complement
 * approximation (MatAssemble_StokesPC_ScaledMassMatrix, stokes_assembly.c:299-450) */
int ptmf_schur_mass_blocks(ptmf_op op, double *Me);

/* ---- measurement --------------------------------------------------------- */
/* Enqueues `reps` back-to-back ptmf_a11_mult (precision 64 or 32) on device
 * arrays and returns the mean time per MatMult from HIP events recorded on the
 * operator's stream around the whole sequence. */
int ptmf_a11_time(ptmf_op op, int precision, const void *x, void *y, int reps,
                  double *ms_per_mult);
/* Same for the element kernel alone (y accumulates; its values are not
 * meaningful afterwards): the duration of the dominant kernel. */
int ptmf_a11_time_kernel(ptmf_op op, int precision, const void *x, void *y, int reps,
                         double *ms_per_kernel);

/* ---- geometric multigrid on A11 (level 0 = coarsest, as PCMG) ---------- */
/* Hierarchy of nlevels boxes, each coarsened by 2 in every direction
 * (pTatin3dStokesBuildMeshHierarchy, ptatin_driver_linear_ts.c:481-513). */
int ptmf_mg_create(int mx, int my, int mz, int nlevels, ptmf_mg *mg);
int ptmf_mg_destroy(ptmf_mg *mg);
int ptmf_mg_default_options(ptmf_mg_options *opt);
int ptmf_mg_levels(ptmf_mg mg, int *nlevels);
/* borrowed operator of a level (do not destroy) */
int ptmf_mg_level_op(ptmf_mg mg, int level, ptmf_op *op);
/* finest coordinates (host, 3N); coarse levels by injection
 * (DMDARestrictCoordinatesHierarchy, dmda_project_coords.c:39-71) */
int ptmf_mg_set_coordinates(ptmf_mg mg, const double *coords);
/* quadrature-point viscosity of one level (host, 27E, element-major) */
int ptmf_mg_set_viscosity(ptmf_mg mg, int level, const double *eta);
/* fill every coarser level from the next finer one, on the device.  Modes
 * 0/1/2: each coarse element takes one arithmetic/geometric/harmonic mean of
 * its 8 children's 216 quadrature values.  Modes 3/4: each coarse quadrature
 * point takes the arithmetic/geometric mean of the children containing it
 * (sub-element variation kept).  Modes 5/6: the reference's coefficient
 * hierarchy from quadrature-point values
 * (SwarmUpdateGaussPropertiesLocalL2Projection_Q1_MPntPStokes_Hierarchy,
 * material_point_utils.c:1138-1265,1267-1420): a nodal field on the finest
 * Q2 lattice (corner nodes weighted by the Q1 basis at each quadrature point,
 * other nodes by 1), carried down by scaled restriction (P^T f)/(P^T 1)
 * (mode 5, -mp_hierarchy_projection_type 0) or injection (mode 6, type 1),
 * and interpolated to each level's quadrature points from the element
 * corners, clamped to the element's nodal range.  The finest level keeps its
 * own values. */
int ptmf_mg_coarsen_viscosity(ptmf_mg mg, int mode);
/* Dirichlet mask of one level (host, 3N bytes) */
int ptmf_mg_set_dirichlet(ptmf_mg mg, int level, const unsigned char *mask);
/* diagonals, Jacobi scalings, Chebyshev eigenvalue bounds, coarse factor */
int ptmf_mg_setup(ptmf_mg mg, const ptmf_mg_options *opt);
/* the coarse factor of the last set-up: kind 1 dense / 2 sparse, its size in
 * bytes, the depth of the nested-dissection tree (0 for dense) and the
 * factorisation time in ms (sparse only; 0 for dense) */
int ptmf_mg_coarse_info(ptmf_mg mg, int *kind, long *bytes, int *depths, double *setup_ms);
int ptmf_mg_level_eigs(ptmf_mg mg, int level, double *emin, double *emax);
int ptmf_mg_set_level_eigs(ptmf_mg mg, int level, double emin, double emax);
/* one V-cycle on the finest level: x = M^-1 b (device FP64 arrays) */
int ptmf_mg_vcycle(ptmf_mg mg, const double *b, double *x);
/* y = A_level x with the level's operator (rediscretised matrix-free or
 * Galerkin stencil), precision 64 (double*) or 32 (float*) */
int ptmf_mg_level_apply(ptmf_mg mg, int level, int precision, const void *x, void *y);
/* The level's operator assembled as a CSR matrix on the host
 * (MatAssemble_StokesA_AUU, stokes_assembly.c:115-297, with the Dirichlet
 * treatment of BCListInsertScaling, dmda_bcs.c:1035-1060: Dirichlet rows and
 * columns removed, unit diagonal inserted), as the reference builds its
 * assembled levels (ptatin_driver_linear_ts.c:640-662).  Rows and columns in
 * the node-interlaced dof order 3n + c, 0-based; a free row holds the
 * couplings to the free dofs of every node sharing an element with its node,
 * columns ascending.  Values from the level's stencil (Galerkin levels) or its
 * MatMult probed with 375 colourings.  First call with colidx = vals = NULL:
 * *nnz, and rowptr (3N + 1 entries) when not NULL; then with colidx and vals
 * of *nnz entries and the same rowptr.  Needs ptmf_mg_setup; host work and
 * PCIe traffic grow as 1125 N doubles, meant for coarse levels. */
int ptmf_mg_level_assemble(ptmf_mg mg, int level, long *nnz, long *rowptr, long *colidx, double *vals);
/* The six rigid-body modes of a level from its node coordinates, the near
 * null space the reference attaches to its assembled levels
 * (ptatin_driver_linear_ts.c:663-666, PETSc's MatNullSpaceCreateRigidBody):
 * modes[k * 3N + 3n + c], k = 0..2 translations along x, y, z, k = 3 rotation
 * about z (-y, x, 0), k = 4 about x (0, -z, y), k = 5 about y (z, 0, -x).
 * Host array of 6 * 3N doubles, not normalised. */
int ptmf_mg_level_rigid_body_modes(ptmf_mg mg, int level, double *modes);
/* yf += P xc and yc = P^T xf between levels level-1 and level (device FP64) */
int ptmf_mg_prolong_add(ptmf_mg mg, int level, const double *xc, double *yf);
int ptmf_mg_restrict(ptmf_mg mg, int level, const double *xf, double *yc);
/* Chebyshev-Jacobi smoothing on one level (precision 64: double*, 32: float*) */
int ptmf_mg_smooth(ptmf_mg mg, int level, int precision, int its, int nonzero_guess,
                   const void *b, void *x);
/* FGMRES(restart) in FP64 on the finest A11 with the V-cycle as right
 * preconditioner (-fieldsplit_u_ksp_type fgmres); x holds the initial guess.
 * history (optional) receives the true residual norm at each restart and the
 * FGMRES estimate after each iteration. */
int ptmf_mg_fgmres(ptmf_mg mg, const double *b, double *x, double rtol, int maxit, int restart,
                   int *its, double *rnorm, double *history, int history_len);
/* Flexible CG preconditioned by the V-cycle, from the given x (a PCG-type
 * outer iteration for the symmetric A11 block; no restart, 5 work vectors).
 * Same stopping rule and outputs as ptmf_mg_fgmres. */
int ptmf_mg_fcg(ptmf_mg mg, const double *b, double *x, double rtol, int maxit, int *its,
                double *rnorm, double *history, int history_len);

/* ---- batched MatMult: many independent boxes, one launch per stage ----- *
 * Batches set-up operators of any sizes (each keeps its own coefficients,
 * tiling and Dirichlet data); ptmf_batch_mult applies MatMult_MFStokes_A11 to
 * every box, x and y being host arrays of nops device pointers (double** for
 * precision 64, float** for 32).  The work of all boxes is one list of
 * (box, tile) items, so small boxes no longer cost a launch each.  No
 * counterpart in the reference (one PETSc Mat per problem). */
int ptmf_batch_create(const ptmf_op *ops, int nops, ptmf_batch *batch);
int ptmf_batch_destroy(ptmf_batch *batch);
int ptmf_batch_mult(ptmf_batch batch, int precision, const void *const *x, void *const *y);
/* ms per batched MatMult over reps back-to-back launches (HIP events) */
int ptmf_batch_time(ptmf_batch batch, int precision, const void *const *x, void *const *y, int reps,
                    double *ms_per_mult);

/* ---- batched multigrid: many independent boxes in lockstep --------------- *
 * No counterpart in the reference (one solve per PETSc communicator there):
 * the velocity-block solves of many independent boxes, each a set-up ptmf_mg
 * (any box sizes; same number of levels and smoother options; lattice
 * transfers, rediscretised levels, point-Jacobi Chebyshev, dense coarse
 * inverse), run with one launch per stage over all boxes and the V-cycle
 * captured once as a HIP graph.  The ptmf_mg objects must outlive the batch
 * and must not be used while it runs (their level buffers are shared). */
int ptmf_bmg_create(const ptmf_mg *mgs, int nbox, ptmf_bmg *out);
int ptmf_bmg_destroy(ptmf_bmg *bmg);
/* x[b] = V-cycle(b[b]) for every box (host arrays of device pointers). */
int ptmf_bmg_vcycle(ptmf_bmg bmg, const double *const *b, double *const *x);
/* Flexible CG (FP64) on every box from the given x[b], preconditioned by the
 * batched V-cycle; box b stops once ||r_b|| <= rtol ||b_b||.  its[b] and
 * rnorm[b] (the true residual at return) are host arrays of nbox entries. */
int ptmf_bmg_fcg(ptmf_bmg bmg, const double *const *b, double *const *x, double rtol, int maxit, int *its,
                 double *rnorm);

/* ---- saddle point [A11 A12; A21 0] on the finest level ------------------ *
 * Vectors are [u (3N) | p (4E)] device arrays.  The solver is FGMRES with the
 * upper block-triangular preconditioner [A11 A12; 0 S]^-1, S the scaled
 * viscous pressure mass (block diagonal, inverted per element) and A11^-1 one
 * V-cycle of this multigrid (precision per ptmf_mg_options) -- the
 * fieldsplit Schur "upper" configuration of the reference
 * (ptatin_driver_linear_ts.c, -pc_fieldsplit_schur_factorization_type upper).
 * With inner_its > 1, A11^-1 is up to inner_its flexible-CG iterations
 * (V-cycle preconditioned) to relative tolerance inner_rtol, as the
 * reference's inner -fieldsplit_u_ksp_type fgmres -fieldsplit_u_ksp_rtol.
 * ptmf_mg_stokes_setup needs ptmf_mg_setup first. */
int ptmf_mg_stokes_setup(ptmf_mg mg, int inner_its, double inner_rtol);
/* Schur approximation: 0 scaled viscous mass (the reference's, default);
 * 1 element blocks of B^T D^-1 B, D the A11 diagonal (no counterpart in the
 * reference).  Type 1 returns PTMF_ERR_ARG, with the mass blocks kept, when
 * an element has fewer than 4 independent free velocity rows (its block is
 * singular). */
int ptmf_mg_stokes_set_schur(ptmf_mg mg, int type);
int ptmf_mg_stokes_apply(ptmf_mg mg, const double *x, double *y);
int ptmf_mg_stokes_fgmres(ptmf_mg mg, const double *b, double *x, double rtol, int maxit, int restart,
                          int *its, double *rnorm, double *history, int history_len);

#ifdef __cplusplus
}
#endif
#endif
