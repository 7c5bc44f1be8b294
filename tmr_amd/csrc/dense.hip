// Coarsest-level direct solve on the GPU: blocked FP64 Cholesky A = L L^T of
// a dense symmetric matrix and the explicit inverse A^-1 = L^-T L^-1 (the
// coarse solve of every V-cycle is then one GEMV, launch_dense_solve).
// Replaces TACSMg's direct coarse solve (src/TMR_RefinementTools.cpp:83-86).
//
// Matrices are row-major mp x mp with mp a multiple of DB (padding rows and
// columns hold the identity), processed in DB x DB blocks:
//   Cholesky (right-looking): factor block (k,k); solve the panel below it;
//     update the trailing lower blocks;
//   L^-1 = X: X[i][i] = L[i][i]^-1 and, right-looking over k,
//     X[i][j] = -X[i][i] sum_{k=j}^{i-1} L[i][k] X[k][j];
//   A^-1 = X^T X, lower blocks then mirrored, written into the m x m result.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

#include "kernels.h"

namespace tmrx {

constexpr int DB = 64;  // block size
constexpr int DT = 256;  // threads per workgroup

// factor the diagonal block k in LDS; *bad = 1 on a non-positive pivot
__global__ void __launch_bounds__(DT) k_potrf_diag(double *__restrict__ W, int mp, int k,
                                                   int *bad) {
  __shared__ double a[DB][DB + 1];
  const int tid = threadIdx.x;
  double *blk = W + (size_t)(k * DB) * mp + k * DB;
  for (int t = tid; t < DB * DB; t += DT) a[t / DB][t % DB] = blk[(size_t)(t / DB) * mp + t % DB];
  __syncthreads();
  for (int j = 0; j < DB; j++) {
    const double d = a[j][j];
    __syncthreads();
    if (!(d > 0.0)) {
      if (tid == 0) atomicOr(bad, 1);
      return;
    }
    const double s = sqrt(d);
    if (tid == 0) a[j][j] = s;
    for (int i = j + 1 + tid; i < DB; i += DT) a[i][j] /= s;
    __syncthreads();
    // trailing update of the lower part: a[i][l] -= a[i][j] a[l][j], j < l <= i
    // (thread: row i = tid / 4, columns l = j + 1 + tid % 4, step 4)
    {
      const int i = tid >> 2;
      if (i > j) {
        const double aij = a[i][j];
        for (int l = j + 1 + (tid & 3); l <= i; l += 4) a[i][l] -= aij * a[l][j];
      }
    }
    __syncthreads();
  }
  for (int t = tid; t < DB * DB; t += DT) {
    const int i = t / DB, l = t % DB;
    blk[(size_t)i * mp + l] = l <= i ? a[i][l] : 0.0;
  }
}

// rows below block k: x L_kk^T = a, one row per thread
__global__ void __launch_bounds__(DT) k_trsm_panel(double *__restrict__ W, int mp, int k) {
  __shared__ double L[DB][DB + 1];
  const int tid = threadIdx.x;
  const double *lk = W + (size_t)(k * DB) * mp + k * DB;
  for (int t = tid; t < DB * DB; t += DT) L[t / DB][t % DB] = lk[(size_t)(t / DB) * mp + t % DB];
  __syncthreads();
  const int r = (k + 1) * DB + blockIdx.x * DT + tid;
  if (r >= mp) return;
  double *row = W + (size_t)r * mp + k * DB;
  double x[DB];
#pragma unroll
  for (int j = 0; j < DB; j++) x[j] = row[j];
#pragma unroll
  for (int j = 0; j < DB; j++) {
    double v = x[j];
#pragma unroll
    for (int t = 0; t < j; t++) v -= x[t] * L[j][t];
    x[j] = v / L[j][j];
  }
#pragma unroll
  for (int j = 0; j < DB; j++) row[j] = x[j];
}

// C[I][J] -= A[I][k] A[J][k]^T for the trailing lower blocks I >= J > k
__global__ void __launch_bounds__(DT) k_syrk_trailing(double *__restrict__ W, int mp, int k) {
  __shared__ double PI[DB][DB + 1], PJ[DB][DB + 1];
  // lower-triangle tile index -> (I, J) over the T = nblk - k - 1 trailing blocks
  const int t = blockIdx.x;
  int I = (int)((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
  while ((I + 1) * (I + 2) / 2 <= t) I++;
  while (I * (I + 1) / 2 > t) I--;
  const int J = t - I * (I + 1) / 2;
  const int bi = (k + 1 + I) * DB, bj = (k + 1 + J) * DB, bk = k * DB;
  const int tid = threadIdx.x;
  for (int q = tid; q < DB * DB; q += DT) {
    PI[q / DB][q % DB] = W[(size_t)(bi + q / DB) * mp + bk + q % DB];
    PJ[q / DB][q % DB] = W[(size_t)(bj + q / DB) * mp + bk + q % DB];
  }
  __syncthreads();
  // 4 x 4 outputs per thread
  const int r0 = (tid / 16) * 4, c0 = (tid % 16) * 4;
  double acc[4][4] = {};
  for (int kk = 0; kk < DB; kk++) {
    double a[4], b[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      a[u] = PI[r0 + u][kk];
      b[u] = PJ[c0 + u][kk];
    }
#pragma unroll
    for (int u = 0; u < 4; u++)
#pragma unroll
      for (int v = 0; v < 4; v++) acc[u][v] += a[u] * b[v];
  }
#pragma unroll
  for (int u = 0; u < 4; u++)
#pragma unroll
    for (int v = 0; v < 4; v++) W[(size_t)(bi + r0 + u) * mp + bj + c0 + v] -= acc[u][v];
}

// X[i][i] = L[i][i]^-1 (lower) for every diagonal block i = blockIdx.x:
// column c solves L x = e_c
__global__ void __launch_bounds__(DB) k_trtri_diag(const double *__restrict__ W,
                                                   double *__restrict__ X, int mp) {
  __shared__ double L[DB][DB + 1];
  const int i = blockIdx.x;
  const int tid = threadIdx.x;
  const size_t o = (size_t)(i * DB) * mp + i * DB;
  for (int t = tid; t < DB * DB; t += DB) L[t / DB][t % DB] = W[o + (size_t)(t / DB) * mp + t % DB];
  __syncthreads();
  const int c = tid;
  double x[DB];
#pragma unroll
  for (int r = 0; r < DB; r++) {
    double v = r == c ? 1.0 : 0.0;
#pragma unroll
    for (int t = 0; t < r; t++) v -= L[r][t] * x[t];
    x[r] = r < c ? 0.0 : v / L[r][r];
  }
#pragma unroll
  for (int r = 0; r < DB; r++) X[o + (size_t)r * mp + c] = x[r];
}

// zero the strictly lower blocks of X (the accumulators of the right-looking
// inverse below)
__global__ void k_trtri_zero(int mp, double *__restrict__ X) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)mp * mp) return;
  const int r = (int)(t / mp), c = (int)(t % mp);
  if (r / DB > c / DB) X[t] = 0.0;
}

// Right-looking L^-1, step k: every block (i, j), i > k >= j, takes
// X[i][j] -= L[i][k] X[k][j]; row k + 1 has then received all its terms and
// is finished in the same pass, X[k+1][j] = X[k+1][k+1] X[k+1][j].  After
// step k rows 0..k+1 are final, so each step is one launch of
// (nblk - 1 - k) (k + 1) independent block products.
__global__ void __launch_bounds__(DT) k_trtri_step(const double *__restrict__ W,
                                                   double *__restrict__ X, int mp, int k) {
  __shared__ double A[DB][DB + 1], B[DB][DB + 1];
  const int t = blockIdx.x;
  const int i = k + 1 + t / (k + 1), j = t % (k + 1);
  const int tid = threadIdx.x;
  const int r0 = (tid / 16) * 4, c0 = (tid % 16) * 4;
  for (int q = tid; q < DB * DB; q += DT) {
    A[q / DB][q % DB] = W[(size_t)(i * DB + q / DB) * mp + k * DB + q % DB];
    B[q / DB][q % DB] = X[(size_t)(k * DB + q / DB) * mp + j * DB + q % DB];
  }
  __syncthreads();
  double acc[4][4];
#pragma unroll
  for (int u = 0; u < 4; u++)
#pragma unroll
    for (int v = 0; v < 4; v++)
      acc[u][v] = X[(size_t)(i * DB + r0 + u) * mp + j * DB + c0 + v];
  for (int kk = 0; kk < DB; kk++) {
    double a[4], b[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      a[u] = A[r0 + u][kk];
      b[u] = B[kk][c0 + u];
    }
#pragma unroll
    for (int u = 0; u < 4; u++)
#pragma unroll
      for (int v = 0; v < 4; v++) acc[u][v] -= a[u] * b[v];
  }
  if (i != k + 1) {
#pragma unroll
    for (int u = 0; u < 4; u++)
#pragma unroll
      for (int v = 0; v < 4; v++) X[(size_t)(i * DB + r0 + u) * mp + j * DB + c0 + v] = acc[u][v];
    return;
  }
  // the last term of row k + 1: X[k+1][j] = X[k+1][k+1] acc
  __syncthreads();
#pragma unroll
  for (int u = 0; u < 4; u++)
#pragma unroll
    for (int v = 0; v < 4; v++) B[r0 + u][c0 + v] = acc[u][v];
  for (int q = tid; q < DB * DB; q += DT)
    A[q / DB][q % DB] = X[(size_t)(i * DB + q / DB) * mp + i * DB + q % DB];
  __syncthreads();
  double out[4][4] = {};
  for (int kk = 0; kk < DB; kk++) {
    double a[4], b[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      a[u] = A[r0 + u][kk];
      b[u] = B[kk][c0 + u];
    }
#pragma unroll
    for (int u = 0; u < 4; u++)
#pragma unroll
      for (int v = 0; v < 4; v++) out[u][v] += a[u] * b[v];
  }
#pragma unroll
  for (int u = 0; u < 4; u++)
#pragma unroll
    for (int v = 0; v < 4; v++) X[(size_t)(i * DB + r0 + u) * mp + j * DB + c0 + v] = out[u][v];
}

// A^-1[I][J] = sum_{k >= I} X[k][I]^T X[k][J] for I >= J (X lower); the m x m
// result (row-major) takes both the block and its mirror
__global__ void __launch_bounds__(DT) k_gram(const double *__restrict__ X, int mp, int nblk,
                                             int m, double *__restrict__ out) {
  __shared__ double A[DB][DB + 1], B[DB][DB + 1];
  const int t = blockIdx.x;
  int I = (int)((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
  while ((I + 1) * (I + 2) / 2 <= t) I++;
  while (I * (I + 1) / 2 > t) I--;
  const int J = t - I * (I + 1) / 2;
  const int tid = threadIdx.x;
  const int r0 = (tid / 16) * 4, c0 = (tid % 16) * 4;
  double acc[4][4] = {};
  for (int k = I; k < nblk; k++) {
    __syncthreads();
    for (int q = tid; q < DB * DB; q += DT) {
      A[q / DB][q % DB] = X[(size_t)(k * DB + q / DB) * mp + I * DB + q % DB];
      B[q / DB][q % DB] = X[(size_t)(k * DB + q / DB) * mp + J * DB + q % DB];
    }
    __syncthreads();
    for (int kk = 0; kk < DB; kk++) {
      double a[4], b[4];
#pragma unroll
      for (int u = 0; u < 4; u++) {
        a[u] = A[kk][r0 + u];
        b[u] = B[kk][c0 + u];
      }
#pragma unroll
      for (int u = 0; u < 4; u++)
#pragma unroll
        for (int v = 0; v < 4; v++) acc[u][v] += a[u] * b[v];
    }
  }
#pragma unroll
  for (int u = 0; u < 4; u++)
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const int r = I * DB + r0 + u, c = J * DB + c0 + v;
      if (r < m && c < m) {
        out[(size_t)r * m + c] = acc[u][v];
        out[(size_t)c * m + r] = acc[u][v];
      }
    }
}

// zero the dense matrix, identity on the padding rows (r >= m)
__global__ void k_dense_fill(int64_t total, int mp, int m, double *__restrict__ W) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= total) return;
  const int r = (int)(t / mp), c = (int)(t % mp);
  W[t] = (r >= m && r == c) ? 1.0 : 0.0;
}

// one thread per block of the component's block rows
__global__ void k_bsr_scatter(const int *__restrict__ col, const double *__restrict__ val, int n0,
                              int k0, int k1, const int *__restrict__ rowof, int mp,
                              double *__restrict__ W) {
  const int k = k0 + blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= k1) return;
  const int i = rowof[k] - n0, j = col[k] - n0;
  const double *v = val + 9 * (size_t)k;
  for (int p = 0; p < 3; p++)
    for (int q = 0; q < 3; q++) W[(size_t)(3 * i + p) * mp + 3 * j + q] = v[3 * p + q];
}

// ---------------------------------------------------------------------------
// Large matrices (m >= DENSE_BIG): the same factorisation by recursive
// halving at multiples of 256, with 256 x 256 diagonal blocks factored and
// inverted by the kernels above (as matrices of their own) and everything
// else in plain FP64 GEMMs (rocBLAS), large enough to fill the device:
//   cholinv(A) = (L, X = L^-1): for A = [A11 .; A21 A22],
//     (L11, X11) = cholinv(A11); L21 = A21 X11^T; A22 -= L21 L21^T;
//     (L22, X22) = cholinv(A22); X21 = -X22 (L21 X11);
//   gram(c) = the lower part of X[c:, c:]^T X[c:, c:] (X lower), by halving
//     the block columns: G22 = gram(second half), G21 = X22^T X21 over all
//     rows below, G11 = gram(first half).
// The factor and the inverse live in W and X; A21 is parked in the free
// block of X where X21 goes, and L21 X11 (transposed) in the unused upper
// block of W, so no GEMM works in place.
constexpr int DS = 256;          // diagonal block of the recursion
constexpr int DENSE_BIG = 2048;  // matrices from this size take this path
constexpr int GRAM_LEAF = 1024;  // gram: blocks up to this size in one GEMM

// out (m x m) = the lower part of G (mp x mp, row-major) and its mirror
__global__ void k_dense_mirror(const double *__restrict__ G, int mp, int m, double *__restrict__ out) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)m * m) return;
  const int r = (int)(t / m), c = (int)(t % m);
  out[t] = r >= c ? G[(size_t)r * mp + c] : G[(size_t)c * mp + r];
}

namespace {

// row-major C (M x N) = alpha op(A) op(B) + beta C on a column-major GEMM:
// C^T = op(B)^T op(A)^T
bool gemm_rm(rocblas_handle h, bool ta, bool tb, int M, int N, int K, double alpha,
             const double *A, int lda, const double *B, int ldb, double beta, double *C,
             int ldc) {
  if (M <= 0 || N <= 0) return true;
  return rocblas_dgemm(h, tb ? rocblas_operation_transpose : rocblas_operation_none,
                       ta ? rocblas_operation_transpose : rocblas_operation_none, N, M, K, &alpha,
                       B, ldb, A, lda, &beta, C, ldc) == rocblas_status_success;
}

bool copy2d(double *dst, int ldd, const double *src, int lds, int rows, int cols, hipStream_t st) {
  return hipMemcpy2DAsync(dst, sizeof(double) * ldd, src, sizeof(double) * lds,
                          sizeof(double) * cols, rows, hipMemcpyDeviceToDevice, st) == hipSuccess;
}

struct Dense {
  rocblas_handle h;
  hipStream_t st;
  double *W, *X, *D, *Xs;  // D, Xs: DS x DS work of the diagonal blocks
  int ld;
  int *bad;
  double *at(double *M, int r, int c) const { return M + (size_t)r * ld + c; }

  // L (lower) and L^-1 of the DS x DS block at (o, o)
  bool leaf(int o) {
    const int nblk = DS / DB;
    if (!copy2d(D, DS, at(W, o, o), ld, DS, DS, st)) return false;
    for (int k = 0; k < nblk; k++) {
      hipLaunchKernelGGL(k_potrf_diag, dim3(1), dim3(DT), 0, st, D, DS, k, bad);
      const int rows = DS - (k + 1) * DB;
      if (rows > 0) {
        hipLaunchKernelGGL(k_trsm_panel, dim3((rows + DT - 1) / DT), dim3(DT), 0, st, D, DS, k);
        const int T = nblk - k - 1;
        hipLaunchKernelGGL(k_syrk_trailing, dim3(T * (T + 1) / 2), dim3(DT), 0, st, D, DS, k);
      }
    }
    if (!copy2d(at(W, o, o), ld, D, DS, DS, DS, st)) return false;
    if (hipMemsetAsync(Xs, 0, sizeof(double) * DS * DS, st) != hipSuccess) return false;
    hipLaunchKernelGGL(k_trtri_diag, dim3(nblk), dim3(DB), 0, st, D, Xs, DS);
    for (int k = 0; k + 1 < nblk; k++)
      hipLaunchKernelGGL(k_trtri_step, dim3((nblk - 1 - k) * (k + 1)), dim3(DT), 0, st, D, Xs, DS,
                         k);
    return copy2d(at(X, o, o), ld, Xs, DS, DS, DS, st);
  }

  // L and X = L^-1 of the n x n diagonal block at (o, o), n a multiple of DS
  bool cholinv(int o, int n) {
    if (n == DS) return leaf(o);
    const int n1 = DS * ((n / DS) / 2), n2 = n - n1, o2 = o + n1;
    if (!cholinv(o, n1)) return false;
    // L21 = A21 X11^T (A21 parked in X21's place)
    if (!copy2d(at(X, o2, o), ld, at(W, o2, o), ld, n2, n1, st)) return false;
    if (!gemm_rm(h, false, true, n2, n1, n1, 1.0, at(X, o2, o), ld, at(X, o, o), ld, 0.0,
                 at(W, o2, o), ld))
      return false;
    // A22 -= L21 L21^T, lower part, in up to four strips of block rows
    const int hs = DS * (((n2 / DS) + 3) / 4);
    for (int s0 = o2; s0 < o2 + n2; s0 += hs) {
      const int s1 = s0 + hs < o2 + n2 ? s0 + hs : o2 + n2;
      if (!gemm_rm(h, false, true, s1 - s0, s1 - o2, n1, -1.0, at(W, s0, o), ld, at(W, o2, o), ld,
                   1.0, at(W, s0, o2), ld))
        return false;
    }
    if (!cholinv(o2, n2)) return false;
    // X21 = -X22 (L21 X11): (L21 X11)^T = X11^T L21^T into the upper block
    if (!gemm_rm(h, true, true, n1, n2, n1, 1.0, at(X, o, o), ld, at(W, o2, o), ld, 0.0,
                 at(W, o, o2), ld))
      return false;
    return gemm_rm(h, false, true, n2, n1, n2, -1.0, at(X, o2, o2), ld, at(W, o, o2), ld, 0.0,
                   at(X, o2, o), ld);
  }

  // lower part of X[o:end, o:o+n]^T X[o:end, o:o+n] into W[o:o+n, o:o+n]
  bool gram(int o, int n, int end) {
    if (n <= GRAM_LEAF)
      return gemm_rm(h, true, false, n, n, end - o, 1.0, at(X, o, o), ld, at(X, o, o), ld, 0.0,
                     at(W, o, o), ld);
    const int n1 = DS * ((n / DS) / 2), n2 = n - n1, o2 = o + n1;
    if (!gram(o2, n2, end)) return false;
    if (!gemm_rm(h, true, false, n2, n1, end - o2, 1.0, at(X, o2, o2), ld, at(X, o2, o), ld, 0.0,
                 at(W, o2, o), ld))
      return false;
    return gram(o, n1, end);
  }
};

// workspace: D (DS^2), Xs (DS^2)
bool dense_inverse_big(rocblas_handle h, double *W, double *X, double *ws, int m, double *out,
                       int *bad, hipStream_t st) {
  const int mp = dense_padded(m);
  if (rocblas_set_stream(h, st) != rocblas_status_success) return false;
  if (hipMemsetAsync(X, 0, sizeof(double) * mp * mp, st) != hipSuccess) return false;
  Dense dn{h, st, W, X, ws, ws + DS * DS, mp, bad};
  if (!dn.cholinv(0, mp)) return false;
  // the factor is no longer needed: A^-1 = X^T X into W
  if (!dn.gram(0, mp, mp)) return false;
  const int64_t tot = (int64_t)m * m;
  hipLaunchKernelGGL(k_dense_mirror, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, W, mp,
                     m, out);
  return true;
}

}  // namespace

void *dense_blas_create() {
  rocblas_handle h = nullptr;
  if (rocblas_create_handle(&h) != rocblas_status_success) return nullptr;
  return h;
}

void dense_blas_destroy(void *h) {
  if (h) (void)rocblas_destroy_handle(static_cast<rocblas_handle>(h));
}

int dense_padded(int m) {
  return m >= DENSE_BIG ? (m + DS - 1) / DS * DS : (m + DB - 1) / DB * DB;
}

size_t dense_workspace(int m) { return m < DENSE_BIG ? 0 : 2 * (size_t)DS * DS; }

int launch_dense_inverse_big(void *blas, double *W, double *X, double *ws, int m, double *out,
                             int *bad, hipStream_t st) {
  if (m < DENSE_BIG || !blas || !ws) return -1;
  if (!dense_inverse_big(static_cast<rocblas_handle>(blas), W, X, ws, m, out, bad, st)) return -1;
  TMRX_LAUNCH_CHECK();
  return 0;
}

void launch_dense_inverse(double *W, double *X, int m, double *out, int *bad, hipStream_t st) {
  const int mp = dense_padded(m), nblk = mp / DB;
  if (m == 0) return;
  for (int k = 0; k < nblk; k++) {
    hipLaunchKernelGGL(k_potrf_diag, dim3(1), dim3(DT), 0, st, W, mp, k, bad);
    const int rows = mp - (k + 1) * DB;
    if (rows > 0) {
      hipLaunchKernelGGL(k_trsm_panel, dim3((rows + DT - 1) / DT), dim3(DT), 0, st, W, mp, k);
      const int T = nblk - k - 1;
      hipLaunchKernelGGL(k_syrk_trailing, dim3(T * (T + 1) / 2), dim3(DT), 0, st, W, mp, k);
    }
  }
  // L^-1: the diagonal blocks at once, then nblk - 1 right-looking steps
  const int64_t tot = (int64_t)mp * mp;
  hipLaunchKernelGGL(k_trtri_zero, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, mp, X);
  hipLaunchKernelGGL(k_trtri_diag, dim3(nblk), dim3(DB), 0, st, W, X, mp);
  for (int k = 0; k + 1 < nblk; k++)
    hipLaunchKernelGGL(k_trtri_step, dim3((nblk - 1 - k) * (k + 1)), dim3(DT), 0, st, W, X, mp, k);
  hipLaunchKernelGGL(k_gram, dim3(nblk * (nblk + 1) / 2), dim3(DT), 0, st, X, mp, nblk, m, out);
  TMRX_LAUNCH_CHECK();
}

void launch_bsr_to_dense(const int *brow, const int *col, const double *val, int n0, int n1,
                         int k0, int k1, int mp, double *W, hipStream_t st) {
  const int m = 3 * (n1 - n0);
  const int64_t total = (int64_t)mp * mp;
  hipLaunchKernelGGL(k_dense_fill, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, total,
                     mp, m, W);
  if (k1 > k0)
    hipLaunchKernelGGL(k_bsr_scatter, dim3((k1 - k0 + 255) / 256), dim3(256), 0, st, col, val, n0,
                       k0, k1, brow, mp, W);
  TMRX_LAUNCH_CHECK();
}

}  // namespace tmrx
