// Design filters on the filter forest's nodal field:
//   TMRHelmholtzFilter (src/topology/TMRHelmholtzFilter.cpp:31-387):
//     x_f = (M + r^2 K)^-1 M x, transpose x = M (M + r^2 K)^-1 g, with M the
//     consistent mass and K the Laplacian of the trilinear scalar element
//     (TMRHexaHelmholtzModel, src/topology/TMRHelmholtzModel.cpp:85-99);
//   TMRMatrixFilter (src/topology/TMRMatrixFilter.cpp:50-368): with
//     D = M 1, a = 1 / (1 + r^2 / D^(2/3)), b = a r^2 / (D^(2/3) D), the
//     N-term Horner sum out = t + b o (M out) from out = t = a o x, scaled
//     by 1 / filter(1).
// M and K are assembled once on the host over the independent nodes (hanging
// nodes through their constraints) as scalar CSR; the products, the
// Horner sums and the conjugate-gradient solve of the Helmholtz operator
// (Jacobi preconditioned, to 1e-12 as the reference's GMRES) run on the GPU.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "../../include/tmrx.h"
#include "filter.h"

namespace tmrx {

namespace {

// trilinear shape functions and gradients on a box of extents h at the
// reference point p in [-1, 1]^3 (corner a: bit d = +1 side of axis d)
void shape(const double h[3], const double p[3], double N[8], double dN[8][3]) {
  for (int a = 0; a < 8; a++) {
    double n[3], dn[3];
    for (int d = 0; d < 3; d++) {
      const double s = ((a >> d) & 1) ? 1.0 : -1.0;
      n[d] = 0.5 * (1.0 + s * p[d]);
      dn[d] = 0.5 * s * 2.0 / h[d];  // d/dx = (2 / h) d/dxi
    }
    N[a] = n[0] * n[1] * n[2];
    dN[a][0] = dn[0] * n[1] * n[2];
    dN[a][1] = n[0] * dn[1] * n[2];
    dN[a][2] = n[0] * n[1] * dn[2];
  }
}

// element mass and Laplacian by 2 x 2 x 2 Gauss quadrature (exact for boxes)
void elem_mats(const double h[3], double Me[64], double Ke[64]) {
  const double g = 1.0 / std::sqrt(3.0);
  const double detJ = h[0] * h[1] * h[2] / 8.0;
  std::fill(Me, Me + 64, 0.0);
  std::fill(Ke, Ke + 64, 0.0);
  for (int q = 0; q < 8; q++) {
    const double p[3] = {(q & 1) ? g : -g, (q & 2) ? g : -g, (q & 4) ? g : -g};
    double N[8], dN[8][3];
    shape(h, p, N, dN);
    for (int a = 0; a < 8; a++)
      for (int b = 0; b < 8; b++) {
        Me[8 * a + b] += detJ * N[a] * N[b];
        Ke[8 * a + b] += detJ * (dN[a][0] * dN[b][0] + dN[a][1] * dN[b][1] + dN[a][2] * dN[b][2]);
      }
  }
}

}  // namespace

int Filter::init(Forest *f, int kind, double r, int nterms) {
  if (!f || (kind != 0 && kind != 1) || !(r > 0.0) || (kind == 1 && nterms < 1)) {
    last_error = "filter: invalid arguments";
    return -1;
  }
  if (!f->blocks_are_boxes()) {
    last_error = "filter: every block must be an axis-aligned box with points set";
    return -1;
  }
  if (f->create_nodes()) {
    last_error = "filter: " + f->last_error;
    return -1;
  }
  kind_ = kind;
  r_ = r;
  nterms_ = nterms;
  n_ = f->num_owned_nodes();
  const auto &octs = f->octants();
  const auto &conn = f->conn();
  const auto &dp = f->dep_ptr();
  const auto &dc = f->dep_conn();
  const auto &dw = f->dep_weights();
  // element contributions through the hanging constraints, as (row, col,
  // mass, laplacian) triplets per row
  std::vector<std::vector<std::pair<int, std::pair<double, double>>>> rows(n_);
  double Me[64], Ke[64];
  for (size_t e = 0; e < octs.size(); e++) {
    double h[3];
    f->element_extent(octs[e], h);
    elem_mats(h, Me, Ke);
    std::vector<std::pair<int, double>> ex[8];
    for (int a = 0; a < 8; a++) {
      const int g = conn[8 * e + a];
      if (g >= 0)
        ex[a].push_back({g, 1.0});
      else
        for (int j = dp[-g - 1]; j < dp[-g]; j++) ex[a].push_back({dc[j], dw[j]});
    }
    for (int a = 0; a < 8; a++)
      for (auto &pa : ex[a])
        for (int b = 0; b < 8; b++)
          for (auto &pb : ex[b]) {
            const double w = pa.second * pb.second;
            rows[pa.first].push_back({pb.first, {w * Me[8 * a + b], w * Ke[8 * a + b]}});
          }
  }
  std::vector<int> ptr(n_ + 1, 0), col;
  std::vector<double> mv, av;
  for (int i = 0; i < n_; i++) {
    auto &R = rows[i];
    std::sort(R.begin(), R.end(), [](const auto &x, const auto &y) { return x.first < y.first; });
    for (size_t k = 0; k < R.size(); k++) {
      if (k == 0 || R[k].first != R[k - 1].first) {
        col.push_back(R[k].first);
        mv.push_back(0.0);
        av.push_back(0.0);
      }
      mv.back() += R[k].second.first;
      av.back() += R[k].second.first + r * r * R[k].second.second;
    }
    ptr[i + 1] = (int)col.size();
    std::vector<std::pair<int, std::pair<double, double>>>().swap(R);
  }
  // Jacobi diagonal of the Helmholtz operator; the matrix filter's scalings
  std::vector<double> dinv(n_), D(n_, 0.0), a(n_), b(n_);
  for (int i = 0; i < n_; i++) {
    for (int k = ptr[i]; k < ptr[i + 1]; k++) {
      if (col[k] == i) dinv[i] = 1.0 / av[k];
      D[i] += mv[k];
    }
    if (D[i] == 0.0) {
      a[i] = 1.0;
      b[i] = 0.0;
    } else {
      const double t = std::pow(D[i], 2.0 / 3.0);
      a[i] = 1.0 / (1.0 + r * r / t);
      b[i] = a[i] * r * r / (t * D[i]);
    }
  }
  if (ptr_.upload(ptr) || col_.upload(col) || mval_.upload(mv) || aval_.upload(av) ||
      dinv_.upload(dinv) || ascale_.upload(a) || bscale_.upload(b) || tinv_.alloc(n_) ||
      w1_.alloc(n_) || w2_.alloc(n_) || w3_.alloc(n_) || w4_.alloc(n_) || w5_.alloc(n_) ||
      in_.alloc(n_) || out_.alloc(n_) || parts_.alloc(filter_reduce_blocks(n_) + 1)) {
    last_error = "filter: device allocation failed";
    return -1;
  }
  if (kind_ == 1) {
    // T^-1 = 1 / filter(1) (with T^-1 = 1 while it is formed)
    launch_filter_fill(n_, 1.0, tinv_.p, st_);
    launch_filter_fill(n_, 1.0, in_.p, st_);
    if (horner(in_.p, out_.p, false)) return -1;
    launch_filter_recip(n_, out_.p, tinv_.p, st_);
    if (hipStreamSynchronize(st_) != hipSuccess) {
      last_error = "filter: device error";
      return -1;
    }
  }
  return 0;
}

// TMRHelmholtzPUFilter (src/topology/TMRHelmholtzPUFilter.cpp:407-700): the
// stencil weights come from the caller (getInteriorStencil /
// getBoundaryStencil); the Horner sum is the matrix filter's with both
// diagonal scalings D^-1 and M = B, and B^T in the transpose.
int Filter::init_custom(int n, int nterms, const int *rowp, const int *cols, const double *B,
                        const double *D) {
  if (n < 1 || nterms < 0 || !rowp || !cols || !B || !D || rowp[0] != 0) {
    last_error = "filter: invalid arguments";
    return -1;
  }
  const int nnz = rowp[n];
  for (int i = 0; i < n; i++) {
    if (rowp[i + 1] < rowp[i]) {
      last_error = "filter: row pointers must not decrease";
      return -1;
    }
    if (!(D[i] > 0.0)) {
      last_error = "filter: the stencil diagonal must be positive";
      return -1;
    }
  }
  for (int k = 0; k < nnz; k++)
    if (cols[k] < 0 || cols[k] >= n) {
      last_error = "filter: column index out of range";
      return -1;
    }
  kind_ = 2;
  nterms_ = nterms;
  n_ = n;
  std::vector<int> ptr(rowp, rowp + n + 1), col(cols, cols + nnz), tptr(n + 1, 0), tcol(nnz);
  std::vector<double> val(B, B + nnz), tval(nnz), dinv(n);
  for (int k = 0; k < nnz; k++) tptr[col[k] + 1]++;
  for (int i = 0; i < n; i++) tptr[i + 1] += tptr[i];
  {
    std::vector<int> fill(tptr.begin(), tptr.end() - 1);
    for (int i = 0; i < n; i++)
      for (int k = ptr[i]; k < ptr[i + 1]; k++) {
        tcol[fill[col[k]]] = i;
        tval[fill[col[k]]++] = val[k];
      }
  }
  for (int i = 0; i < n; i++) dinv[i] = 1.0 / D[i];
  if (nnz == 0) {
    col.push_back(0);
    val.push_back(0.0);
    tcol.push_back(0);
    tval.push_back(0.0);
  }
  if (ptr_.upload(ptr) || col_.upload(col) || mval_.upload(val) || tptr_.upload(tptr) ||
      tcol_.upload(tcol) || tval_.upload(tval) || ascale_.upload(dinv) || bscale_.upload(dinv) ||
      tinv_.alloc(n_) || w1_.alloc(n_) || w2_.alloc(n_) || w5_.alloc(n_) || in_.alloc(n_) ||
      out_.alloc(n_)) {
    last_error = "filter: device allocation failed";
    return -1;
  }
  // T^-1 = 1 / filter(1), formed with T^-1 = 1
  launch_filter_fill(n_, 1.0, tinv_.p, st_);
  launch_filter_fill(n_, 1.0, in_.p, st_);
  if (horner(in_.p, out_.p, false)) return -1;
  launch_filter_recip(n_, out_.p, tinv_.p, st_);
  if (hipStreamSynchronize(st_) != hipSuccess) {
    last_error = "filter: device error";
    return -1;
  }
  return 0;
}

// Horner sum of the matrix filter (transpose: the same sum with M^T and the
// diagonal scalings in the opposite order)
int Filter::horner(const double *x, double *y, bool transpose) {
  double *t = w1_.p, *u = w2_.p;
  if (!transpose) {
    launch_filter_mul(n_, ascale_.p, x, t, st_);  // t = a o x
    launch_filter_copy(n_, t, y, st_);
    for (int k = 0; k < nterms_; k++) {
      launch_filter_spmv(n_, ptr_.p, col_.p, mval_.p, y, u, st_);  // u = M y
      launch_filter_mul_add(n_, bscale_.p, u, t, y, st_);          // y = b o u + t
    }
    launch_filter_mul(n_, tinv_.p, y, y, st_);
  } else {
    launch_filter_mul(n_, tinv_.p, x, t, st_);  // t = T^-1 o x
    launch_filter_copy(n_, t, y, st_);
    const bool sym = kind_ != 2;  // the matrix filter's M is symmetric
    for (int k = 0; k < nterms_; k++) {
      launch_filter_mul(n_, bscale_.p, y, u, st_);  // u = b o y
      launch_filter_spmv(n_, sym ? ptr_.p : tptr_.p, sym ? col_.p : tcol_.p,
                         sym ? mval_.p : tval_.p, u, y, st_);  // y = M^T u
      launch_filter_add(n_, t, y, st_);                        // y += t
    }
    launch_filter_mul(n_, ascale_.p, y, y, st_);
  }
  return 0;
}

double Filter::dot(const double *a, const double *b) {
  const int nb = filter_reduce_blocks(n_);
  launch_filter_dot(n_, a, b, parts_.p, st_);
  std::vector<double> h(nb);
  if (hipMemcpyAsync(h.data(), parts_.p, nb * sizeof(double), hipMemcpyDeviceToHost, st_) !=
          hipSuccess ||
      hipStreamSynchronize(st_) != hipSuccess)
    return std::nan("");
  double s = 0.0;
  for (double v : h) s += v;
  return s;
}

// (M + r^2 K) x = rhs by Jacobi-preconditioned CG from zero to rtol 1e-12
int Filter::helmholtz_solve(const double *rhs, double *x) {
  double *r = w1_.p, *z = w2_.p, *p = w3_.p, *q = w4_.p;
  launch_filter_fill(n_, 0.0, x, st_);
  launch_filter_copy(n_, rhs, r, st_);
  const double bb = dot(rhs, rhs);
  if (!(bb == bb)) {
    last_error = "filter: device error";
    return -1;
  }
  iters_ = 0;
  if (bb == 0.0) return 0;
  launch_filter_mul(n_, dinv_.p, r, z, st_);
  launch_filter_copy(n_, z, p, st_);
  double rz = dot(r, z);
  const double tol2 = 1e-24 * bb;
  for (int it = 1; it <= 2000; it++) {
    launch_filter_spmv(n_, ptr_.p, col_.p, aval_.p, p, q, st_);
    const double alpha = rz / dot(p, q);
    launch_filter_axpy(n_, alpha, p, x, st_);
    launch_filter_axpy(n_, -alpha, q, r, st_);
    const double rr = dot(r, r);
    iters_ = it;
    if (rr <= tol2) return 0;
    launch_filter_mul(n_, dinv_.p, r, z, st_);
    const double rzn = dot(r, z);
    launch_filter_xpby(n_, z, rzn / rz, p, st_);  // p = z + beta p
    rz = rzn;
  }
  last_error = "filter: the Helmholtz solve did not converge";
  return -1;
}

int Filter::apply(int nv, const double *in, double *out, bool on_device, bool transpose) {
  if (nv < 1 || !in || !out) {
    last_error = "filter: invalid arguments";
    return -1;
  }
  const size_t n = (size_t)n_ * nv;
  DevArray<double> din, dout;
  const double *xin = in;
  double *xout = out;
  if (!on_device) {
    if (din.upload(in, n) || dout.alloc(n)) {
      last_error = "filter: device allocation failed";
      return -1;
    }
    xin = din.p;
    xout = dout.p;
  }
  for (int v = 0; v < nv; v++) {
    launch_filter_gather(n_, nv, v, xin, in_.p, st_);
    if (kind_ == 1 || kind_ == 2) {
      if (horner(in_.p, out_.p, transpose)) return -1;
    } else if (!transpose) {
      launch_filter_spmv(n_, ptr_.p, col_.p, mval_.p, in_.p, w5_.p, st_);  // rhs = M x
      if (helmholtz_solve(w5_.p, out_.p)) return -1;
    } else {
      if (helmholtz_solve(in_.p, w5_.p)) return -1;
      launch_filter_spmv(n_, ptr_.p, col_.p, mval_.p, w5_.p, out_.p, st_);  // M psi
    }
    launch_filter_scatter(n_, nv, v, out_.p, xout, st_);
  }
  if (!on_device &&
      hipMemcpyAsync(out, dout.p, n * sizeof(double), hipMemcpyDeviceToHost, st_) != hipSuccess) {
    last_error = "filter: copy failed";
    return -1;
  }
  if (hipStreamSynchronize(st_) != hipSuccess || take_launch_error() != hipSuccess) {
    last_error = "filter: device error";
    return -1;
  }
  return 0;
}

}  // namespace tmrx

// ---------------------------------------------------------------------------
struct tmrx_filter {
  tmrx::Filter *f;
  int device;
};

namespace tmrx {
void set_error(const std::string &s);
}
tmrx::Forest *tmrx_forest_impl(tmrx_forest *f);

extern "C" {

int tmrx_filter_create(tmrx_forest *forest, int kind, double r, int nterms, int device,
                       tmrx_filter **out) {
  if (!out) {
    tmrx::set_error("tmrx_filter_create: null output");
    return -1;
  }
  if (hipSetDevice(device) != hipSuccess) {
    tmrx::set_error("tmrx_filter_create: hipSetDevice failed");
    return -1;
  }
  auto *f = new tmrx::Filter();
  if (f->init(tmrx_forest_impl(forest), kind, r, nterms)) {
    tmrx::set_error(f->last_error);
    delete f;
    return -1;
  }
  *out = new tmrx_filter{f, device};
  return 0;
}

int tmrx_filter_create_custom(int n, int nterms, const int *rowp, const int *cols, const double *B,
                              const double *D, int device, tmrx_filter **out) {
  if (!out) {
    tmrx::set_error("tmrx_filter_create_custom: null output");
    return -1;
  }
  if (hipSetDevice(device) != hipSuccess) {
    tmrx::set_error("tmrx_filter_create_custom: hipSetDevice failed");
    return -1;
  }
  auto *f = new tmrx::Filter();
  if (f->init_custom(n, nterms, rowp, cols, B, D)) {
    tmrx::set_error(f->last_error);
    delete f;
    return -1;
  }
  *out = new tmrx_filter{f, device};
  return 0;
}

void tmrx_filter_free(tmrx_filter *h) {
  if (!h) return;
  (void)hipSetDevice(h->device);
  delete h->f;
  delete h;
}

static int filter_apply(tmrx_filter *h, int nv, const double *in, double *out, int on_device,
                        bool transpose) {
  if (!h || !h->f) {
    tmrx::set_error("null filter handle");
    return -1;
  }
  (void)hipSetDevice(h->device);
  if (h->f->apply(nv, in, out, on_device != 0, transpose)) {
    tmrx::set_error(h->f->last_error);
    return -1;
  }
  return 0;
}

int tmrx_filter_apply(tmrx_filter *h, int vars_per_node, const double *in, double *out,
                      int on_device) {
  return filter_apply(h, vars_per_node, in, out, on_device, false);
}

int tmrx_filter_apply_transpose(tmrx_filter *h, int vars_per_node, const double *in, double *out,
                                int on_device) {
  return filter_apply(h, vars_per_node, in, out, on_device, true);
}

int tmrx_filter_info(tmrx_filter *h, int *num_nodes, int *last_iters) {
  if (!h || !h->f) {
    tmrx::set_error("null filter handle");
    return -1;
  }
  if (num_nodes) *num_nodes = h->f->num_nodes();
  if (last_iters) *last_iters = h->f->last_iters();
  return 0;
}

}  // extern "C"
