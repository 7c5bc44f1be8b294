// Helmholtz and matrix design filters on a filter forest (see filter.cpp).
#pragma once
#include <hip/hip_runtime_api.h>

#include <string>

#include "forest.h"
#include "problem.h"

namespace tmrx {

class Filter {
 public:
  // kind 0: Helmholtz radius r; kind 1: matrix filter, radius r, nterms terms
  int init(Forest *f, int kind, double r, int nterms);
  // kind 2: partition-of-unity filter from given stencils, F = T^-1 sum_k
  // (D^-1 B)^k D^-1 (nterms + 1 terms), B as CSR over n nodes, D > 0
  int init_custom(int n, int nterms, const int *rowp, const int *cols, const double *B,
                  const double *D);
  // out = F in (or F^T in), nv values per independent node interleaved
  int apply(int nv, const double *in, double *out, bool on_device, bool transpose);
  int num_nodes() const { return n_; }
  int last_iters() const { return iters_; }
  std::string last_error;

 private:
  int horner(const double *x, double *y, bool transpose);
  int helmholtz_solve(const double *rhs, double *x);
  double dot(const double *a, const double *b);

  int kind_ = 0, nterms_ = 0, n_ = 0, iters_ = 0;
  double r_ = 0.0;
  hipStream_t st_ = nullptr;
  DevArray<int> ptr_, col_, tptr_, tcol_;  // t*: the transpose (kind 2, B not symmetric)
  DevArray<double> tval_;
  DevArray<double> mval_, aval_, dinv_, ascale_, bscale_, tinv_;
  DevArray<double> w1_, w2_, w3_, w4_, w5_, in_, out_, parts_;
};

// kernels (filter.hip)
int filter_reduce_blocks(int n);
void launch_filter_spmv(int n, const int *ptr, const int *col, const double *val,
                        const double *x, double *y, hipStream_t st);
void launch_filter_fill(int n, double v, double *x, hipStream_t st);
void launch_filter_copy(int n, const double *x, double *y, hipStream_t st);
// y = a o x (elementwise; y may alias x)
void launch_filter_mul(int n, const double *a, const double *x, double *y, hipStream_t st);
// y = a o u + t
void launch_filter_mul_add(int n, const double *a, const double *u, const double *t, double *y,
                           hipStream_t st);
void launch_filter_add(int n, const double *t, double *y, hipStream_t st);
void launch_filter_axpy(int n, double alpha, const double *x, double *y, hipStream_t st);
// p = z + beta p
void launch_filter_xpby(int n, const double *z, double beta, double *p, hipStream_t st);
// y[i] = 1 / x[i] (0 where x is 0)
void launch_filter_recip(int n, const double *x, double *y, hipStream_t st);
// per-block partial sums of a o b
void launch_filter_dot(int n, const double *a, const double *b, double *parts, hipStream_t st);
// component v of an interleaved nv-vector to/from a plain one
void launch_filter_gather(int n, int nv, int v, const double *x, double *y, hipStream_t st);
void launch_filter_scatter(int n, int nv, int v, const double *x, double *y, hipStream_t st);

}  // namespace tmrx
