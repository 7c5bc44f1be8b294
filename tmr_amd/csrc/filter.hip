// Kernels of the design filters (filter.cpp): scalar CSR product, vector
// updates and fixed-order block reductions.
#include <hip/hip_runtime.h>

#include "filter.h"
#include "kernels.h"

namespace tmrx {

constexpr int FB = 256;
static inline dim3 grid_of(int n) { return dim3((unsigned)((n + FB - 1) / FB)); }

__global__ void k_filter_spmv(int n, const int *__restrict__ ptr, const int *__restrict__ col,
                              const double *__restrict__ val, const double *__restrict__ x,
                              double *__restrict__ y) {
  const int i = blockIdx.x * FB + threadIdx.x;
  if (i >= n) return;
  double s = 0.0;
  for (int k = ptr[i]; k < ptr[i + 1]; k++) s += val[k] * x[col[k]];
  y[i] = s;
}
void launch_filter_spmv(int n, const int *ptr, const int *col, const double *val,
                        const double *x, double *y, hipStream_t st) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_filter_spmv, grid_of(n), dim3(FB), 0, st, n, ptr, col, val, x, y);
  TMRX_LAUNCH_CHECK();
}

__global__ void k_filter_fill(int n, double v, double *x) {
  const int i = blockIdx.x * FB + threadIdx.x;
  if (i < n) x[i] = v;
}
void launch_filter_fill(int n, double v, double *x, hipStream_t st) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_filter_fill, grid_of(n), dim3(FB), 0, st, n, v, x);
  TMRX_LAUNCH_CHECK();
}

__global__ void k_filter_lin(int n, int op, double alpha, const double *a, const double *x,
                             const double *t, double *y) {
  const int i = blockIdx.x * FB + threadIdx.x;
  if (i >= n) return;
  switch (op) {
    case 0: y[i] = x[i]; break;                       // copy
    case 1: y[i] = a[i] * x[i]; break;                // elementwise product
    case 2: y[i] = a[i] * x[i] + t[i]; break;         // product plus t
    case 3: y[i] += t[i]; break;                      // add
    case 4: y[i] += alpha * x[i]; break;              // axpy
    case 5: y[i] = x[i] + alpha * y[i]; break;        // x + beta y
    case 6: y[i] = x[i] != 0.0 ? 1.0 / x[i] : 0.0; break;  // reciprocal
  }
}
static void lin(int n, int op, double alpha, const double *a, const double *x, const double *t,
                double *y, hipStream_t st) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_filter_lin, grid_of(n), dim3(FB), 0, st, n, op, alpha, a, x, t, y);
  TMRX_LAUNCH_CHECK();
}
void launch_filter_copy(int n, const double *x, double *y, hipStream_t st) {
  lin(n, 0, 0.0, nullptr, x, nullptr, y, st);
}
void launch_filter_mul(int n, const double *a, const double *x, double *y, hipStream_t st) {
  lin(n, 1, 0.0, a, x, nullptr, y, st);
}
void launch_filter_mul_add(int n, const double *a, const double *u, const double *t, double *y,
                           hipStream_t st) {
  lin(n, 2, 0.0, a, u, t, y, st);
}
void launch_filter_add(int n, const double *t, double *y, hipStream_t st) {
  lin(n, 3, 0.0, nullptr, nullptr, t, y, st);
}
void launch_filter_axpy(int n, double alpha, const double *x, double *y, hipStream_t st) {
  lin(n, 4, alpha, nullptr, x, nullptr, y, st);
}
void launch_filter_xpby(int n, const double *z, double beta, double *p, hipStream_t st) {
  lin(n, 5, beta, nullptr, z, nullptr, p, st);
}
void launch_filter_recip(int n, const double *x, double *y, hipStream_t st) {
  lin(n, 6, 0.0, nullptr, x, nullptr, y, st);
}

int filter_reduce_blocks(int n) { return (n + 4 * FB - 1) / (4 * FB); }

__global__ void k_filter_dot(int n, const double *__restrict__ a, const double *__restrict__ b,
                             double *__restrict__ parts) {
  __shared__ double red[FB];
  double s = 0.0;
  const int base = blockIdx.x * 4 * FB;
  for (int k = 0; k < 4; k++) {
    const int i = base + k * FB + threadIdx.x;
    if (i < n) s += a[i] * b[i];
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = FB / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) parts[blockIdx.x] = red[0];
}
void launch_filter_dot(int n, const double *a, const double *b, double *parts, hipStream_t st) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_filter_dot, dim3(filter_reduce_blocks(n)), dim3(FB), 0, st, n, a, b,
                     parts);
  TMRX_LAUNCH_CHECK();
}

__global__ void k_filter_stride(int n, int nv, int v, int dir, const double *x, double *y) {
  const int i = blockIdx.x * FB + threadIdx.x;
  if (i >= n) return;
  if (dir == 0)
    y[i] = x[(size_t)i * nv + v];
  else
    y[(size_t)i * nv + v] = x[i];
}
void launch_filter_gather(int n, int nv, int v, const double *x, double *y, hipStream_t st) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_filter_stride, grid_of(n), dim3(FB), 0, st, n, nv, v, 0, x, y);
  TMRX_LAUNCH_CHECK();
}
void launch_filter_scatter(int n, int nv, int v, const double *x, double *y, hipStream_t st) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_filter_stride, grid_of(n), dim3(FB), 0, st, n, nv, v, 1, x, y);
  TMRX_LAUNCH_CHECK();
}

}  // namespace tmrx
