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
 * stream; entry points that return host-visible results synchronize it.
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

/* ---- library / device helpers ------------------------------------------ */
const char *ptmf_last_error(void);
int ptmf_version(int *major, int *minor);
int ptmf_device_count(int *n);
int ptmf_set_device(int dev);
int ptmf_malloc(void **dptr, size_t bytes);
int ptmf_free(void *dptr);
int ptmf_memcpy_htod(void *dst, const void *src, size_t bytes);
int ptmf_memcpy_dtoh(void *dst, const void *src, size_t bytes);
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
/* Work decomposition of the MatMult kernel: TX x TY element tiles (TX*TY <= 21)
 * marching through ZC element layers.  A default is chosen at creation; the
 * autotuner overrides it.  No reference counterpart (the CUDA backend's
 * coloring, stokes_operators_cuda.cu:337-397, plays this role there). */
int ptmf_op_set_tiling(ptmf_op op, int TX, int TY, int ZC);
int ptmf_op_get_tiling(ptmf_op op, int *TX, int *TY, int *ZC);
/* coefficient state counter, bumped by every set_* call (replaces the
 * PetscObjectState stamp of stokes_operators.c:1162-1163) */
int ptmf_op_state(ptmf_op op, unsigned long *state);

/* ---- A11 (viscous block) ------------------------------------------------- */
/* y += A11_local u with no boundary handling: the SpMVOp_MatMult contract
 * (stokes_operators.h:59; MFStokesWrapper_A11_Tensor, stokes_operators_tensor.c:181). */
int ptmf_a11_spmv(ptmf_op op, const double *u, double *y);
/* Same contract on HOST arrays (for MFStokesWrapper_A11_HIP in a PETSc
 * adapter): copies u in, adds A u into Yu on the host. */
int ptmf_a11_spmv_host(ptmf_op op, const double *u, double *Yu);
/* MatMult_MFStokes_A11: y = A11 * mask(x) on free rows, y = x on Dirichlet rows. */
int ptmf_a11_mult(ptmf_op op, const double *x, double *y);
/* FP32 version of ptmf_a11_mult (FP32 coordinates, viscosity and accumulation). */
int ptmf_a11_mult_f32(ptmf_op op, const float *x, float *y);
/* MatGetDiagonal_MFStokes_A11: diag(A11), 1 in Dirichlet slots. */
int ptmf_a11_diagonal(ptmf_op op, double *d);

/* ---- B / B^T ------------------------------------------------------------- */
/* MatMult_MFStokes_A12: yu = B p (p: 4E), Dirichlet velocity rows set to 0. */
int ptmf_a12_mult(ptmf_op op, const double *p, double *yu);
/* MatMult_MFStokes_A21: yp = B^T mask(u) (yp: 4E). */
int ptmf_a21_mult(ptmf_op op, const double *u, double *yp);

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

#ifdef __cplusplus
}
#endif
#endif
