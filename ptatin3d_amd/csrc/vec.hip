// vec.hip -- O(N) device kernels: fills and the Dirichlet-row fix-ups of the
// BCList operations (dmda_bcs.c:409-608).
#include "ptmf_internal.h"

namespace ptmf {
namespace {

template <typename T>
__global__ void fill_kernel(T *__restrict__ x, T v, long n) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    x[i] = v;
}

// y[idx[k]] = x[idx[k]]  (BCListInsertDirichlet_MatMult with scale 1)
template <typename T>
__global__ void bc_copy_kernel(const long *__restrict__ idx, long n, const T *__restrict__ x,
                               T *__restrict__ y) {
  for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (long)gridDim.x * blockDim.x) {
    const long g = idx[k];
    y[g] = x[g];
  }
}

// y[idx[k]] = v  (BCListInsertZero / BCListInsertValueIntoDirichletSlot)
template <typename T>
__global__ void bc_set_kernel(const long *__restrict__ idx, long n, T v, T *__restrict__ y) {
  for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (long)gridDim.x * blockDim.x)
    y[idx[k]] = v;
}

// plain 16-byte-vector copy (also the reference kernel for HBM counters)
__global__ void __launch_bounds__(256) copy16_kernel(const uint4 *__restrict__ x, uint4 *__restrict__ y, long n) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    y[i] = x[i];
}

// element-major [27e+q] -> quadrature-point-major [q*nel+e], FP64 and FP32
__global__ void eta_qpmajor_kernel(long nel, const double *__restrict__ em, double *__restrict__ q64,
                                   float *__restrict__ q32) {
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < 27 * nel; t += (long)gridDim.x * blockDim.x) {
    const long q = t / nel, e = t - q * nel;
    const double v = em[27 * e + q];
    q64[t] = v;
    q32[t] = (float)v;
  }
}

inline unsigned grid_for(long n, int bs) {
  long g = (n + bs - 1) / bs;
  if (g > 65536) g = 65536;
  if (g < 1) g = 1;
  return (unsigned)g;
}

}  // namespace

template <typename T>
int launch_fill(T *x, T v, long n, hipStream_t s) {
  if (n <= 0) return PTMF_OK;
  fill_kernel<T><<<grid_for(n, 256), 256, 0, s>>>(x, v, n);
  PTMF_HIP(hipGetLastError());
  return PTMF_OK;
}

template <typename T>
int launch_bc_copy(const long *idx, long n, const T *x, T *y, hipStream_t s) {
  if (n <= 0) return PTMF_OK;
  bc_copy_kernel<T><<<grid_for(n, 256), 256, 0, s>>>(idx, n, x, y);
  PTMF_HIP(hipGetLastError());
  return PTMF_OK;
}

template <typename T>
int launch_bc_set(const long *idx, long n, T v, T *y, hipStream_t s) {
  if (n <= 0) return PTMF_OK;
  bc_set_kernel<T><<<grid_for(n, 256), 256, 0, s>>>(idx, n, v, y);
  PTMF_HIP(hipGetLastError());
  return PTMF_OK;
}

int launch_copy16(const void *x, void *y, long n16, hipStream_t s) {
  if (n16 <= 0) return PTMF_OK;
  copy16_kernel<<<grid_for(n16, 256), 256, 0, s>>>((const uint4 *)x, (uint4 *)y, n16);
  PTMF_HIP(hipGetLastError());
  return PTMF_OK;
}

int launch_eta_qpmajor(long nel, const double *em, double *q64, float *q32, hipStream_t s) {
  eta_qpmajor_kernel<<<grid_for(27 * nel, 256), 256, 0, s>>>(nel, em, q64, q32);
  PTMF_HIP(hipGetLastError());
  return PTMF_OK;
}

template int launch_fill<double>(double *, double, long, hipStream_t);
template int launch_fill<float>(float *, float, long, hipStream_t);
template int launch_fill<int>(int *, int, long, hipStream_t);
template int launch_bc_copy<double>(const long *, long, const double *, double *, hipStream_t);
template int launch_bc_copy<float>(const long *, long, const float *, float *, hipStream_t);
template int launch_bc_set<double>(const long *, long, double, double *, hipStream_t);
template int launch_bc_set<float>(const long *, long, float, float *, hipStream_t);

}  // namespace ptmf
