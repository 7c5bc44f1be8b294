// Galerkin coarse operators for the multigrid hierarchy: TACSMg with
// use_galerkin (TMR_CreateTACSMg, src/TMR_RefinementTools.cpp:37-107) forms
// A_c = P^T A_f P.  Host side: the check that makes the element-wise form
// exact, the BSR structures and term lists, and the per-design setup that
// runs the kernels of galerkin.hip.
#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "problem.h"

namespace tmrx {

#define GCHK(x)                                                          \
  do {                                                                   \
    hipError_t e_ = (x);                                                 \
    if (e_ == hipSuccess) e_ = take_launch_error();                      \
    if (e_ != hipSuccess) {                                              \
      last_error = std::string("HIP error: ") + hipGetErrorString(e_) + \
                   " at " + #x;                                          \
      return -1;                                                         \
    }                                                                    \
  } while (0)

namespace {

// independent nodes and weights of element corner value g (>= 0: node g,
// < 0: hanging node -(g+1) through its constraint)
template <class F>
void expand_node(int g, const std::vector<int> &dptr, const std::vector<int> &dconn,
                 const std::vector<double> &dw, F fn) {
  if (g >= 0) {
    fn(g, 1.0);
  } else {
    for (int j = dptr[-g - 1]; j < dptr[-g]; j++) fn(dconn[j], dw[j]);
  }
}

// trilinear weight of coarse corner c at fine corner a of a child placed at
// `place` (level difference << 24 | z << 16 | y << 8 | x)
double child_weight(int place, int a, int c) {
  const int dl = (place >> 24) & 15;
  const int off[3] = {place & 255, (place >> 8) & 255, (place >> 16) & 255};
  double w = 1.0;
  for (int d = 0; d < 3; d++) {
    const double xi = std::ldexp((double)(off[d] + ((a >> d) & 1)), -dl);
    w *= ((c >> d) & 1) ? xi : 1.0 - xi;
  }
  return w;
}

}  // namespace

template <>
BSR<double> Problem::bsr<double>(int l) {
  Level &L = levels_[l];
  BSR<double> A;
  A.n = L.n_nodes;
  A.nnzb = L.nnzb;
  A.ptr = L.bsr_ptr.p;
  A.col = L.bsr_col.p;
  A.val = L.bsr64.p;
  A.bcmask = L.d_bcmask.p;
  return A;
}
template <>
BSR<float> Problem::bsr<float>(int l) {
  Level &L = levels_[l];
  BSR<float> A;
  A.n = L.n_nodes;
  A.nnzb = L.nnzb;
  A.ptr = L.bsr_ptr.p;
  A.col = L.bsr_col.p;
  A.val = L.bsr32.p;
  A.bcmask = L.d_bcmask.p;
  return A;
}

// The element-wise Galerkin product equals P^T A P only if, on every fine
// element e inside coarse element E, the prolongation of each corner of e
// (a hanging corner through its constraint) is the trilinear interpolation
// of E's corners (each expanded through E's hanging constraints).  Checked
// here once, exactly (the weights are dyadic).
int Problem::build_galerkin() {
  const int nl = (int)levels_.size();
  for (int l = 0; l + 1 < nl; l++) {
    Level &F = levels_[l], &Cc = levels_[l + 1];
    if (!F.nested) {  // the element-wise product needs coarse elements that hold fine ones
      last_error = "galerkin: level " + std::to_string(l + 1) +
                   " is not a coarsen() of the level above; Galerkin coarse operators need nested forests";
      return -1;
    }
    InterpCSR P;
    int dev = 0;
    (void)hipGetDevice(&dev);
    if (F.forest->create_interpolation_device(Cc.forest, P, dev)) {
      last_error = "galerkin: interpolation " + std::to_string(l) + ": " + F.forest->last_error;
      return -1;
    }
    const auto &fdp = F.forest->dep_ptr();
    const auto &fdc = F.forest->dep_conn();
    const auto &fdw = F.forest->dep_weights();
    const auto &cdp = Cc.forest->dep_ptr();
    const auto &cdc = Cc.forest->dep_conn();
    const auto &cdw = Cc.forest->dep_weights();
    std::vector<double> acc(Cc.n_nodes, 0.0);
    std::vector<int> touched;
    auto add = [&](int node, double w) {
      if (acc[node] == 0.0) touched.push_back(node);
      acc[node] += w;
    };
    for (int E = 0; E < Cc.n_elem; E++) {
      for (int k = F.sc_ptr_h[E]; k < F.sc_ptr_h[E + 1]; k++) {
        const int e = F.sc_child_h[k], pl = F.place_h[k];
        for (int a = 0; a < 8; a++) {
          touched.clear();
          // + prolongation of fine corner a
          expand_node(F.conn[8 * e + a], fdp, fdc, fdw, [&](int r, double wr) {
            for (int j = P.rowp[r]; j < P.rowp[r + 1]; j++) add(P.cols[j], wr * P.w[j]);
          });
          // - trilinear interpolation of E's corners
          for (int c = 0; c < 8; c++) {
            const double q = child_weight(pl, a, c);
            if (q == 0.0) continue;
            expand_node(Cc.conn[8 * E + c], cdp, cdc, cdw,
                        [&](int node, double w) { add(node, -q * w); });
          }
          bool ok = true;
          for (int node : touched) {
            if (std::fabs(acc[node]) > 1e-12) ok = false;
            acc[node] = 0.0;
          }
          if (!ok) {
            last_error = "galerkin: the prolongation of level " + std::to_string(l) +
                         " is not the trilinear interpolation inside coarse element " +
                         std::to_string(E);
            return -1;
          }
        }
      }
    }
  }
  for (int l = 1; l < nl; l++) {
    Level &L = levels_[l];
    if (L.KE.alloc(576 * (size_t)L.n_elem)) {
      last_error = "galerkin: device allocation failed";
      return -1;
    }
    if (l + 1 < nl && build_bsr(l)) return -1;
  }
  if (Kgd_.upload(Kgh_)) {
    last_error = "galerkin: device allocation failed";
    return -1;
  }
  galerkin_built_ = true;
  return 0;
}

// BSR pattern of level l and, per block, the list of (element, corner a,
// corner b, weight) terms w K_E[a][b] that sum to it: row i takes corner a
// of every element whose expanded corner a holds i, column j likewise for
// corner b; terms sorted by (column, element, a, b), so the sums run in a
// fixed order.
int Problem::build_bsr(int l) {
  Level &L = levels_[l];
  const auto &dp = L.forest->dep_ptr();
  const auto &dc = L.forest->dep_conn();
  const auto &dw = L.forest->dep_weights();
  const int n = L.n_nodes;
  std::vector<int64_t> rcnt(n + 1, 0);
  for (int E = 0; E < L.n_elem; E++) {
    int nexp = 0;
    for (int b = 0; b < 8; b++) expand_node(L.conn[8 * E + b], dp, dc, dw, [&](int, double) { nexp++; });
    for (int a = 0; a < 8; a++)
      expand_node(L.conn[8 * E + a], dp, dc, dw, [&](int i, double) { rcnt[i + 1] += nexp; });
  }
  for (int i = 0; i < n; i++) rcnt[i + 1] += rcnt[i];
  const int64_t nt = rcnt[n];
  if (nt >= ((int64_t)1 << 31) || 64 * (int64_t)L.n_elem >= ((int64_t)1 << 31)) {
    last_error = "galerkin: level too large for 32-bit term lists";
    return -1;
  }
  struct Term {
    int j, code;
    double w;
  };
  std::vector<Term> terms(nt);
  {
    std::vector<int64_t> fill(rcnt.begin(), rcnt.end() - 1);
    for (int E = 0; E < L.n_elem; E++)
      for (int a = 0; a < 8; a++)
        expand_node(L.conn[8 * E + a], dp, dc, dw, [&](int i, double wa) {
          for (int b = 0; b < 8; b++)
            expand_node(L.conn[8 * E + b], dp, dc, dw, [&](int j, double wb) {
              terms[fill[i]++] = Term{j, 64 * E + 8 * a + b, wa * wb};
            });
        });
  }
  std::vector<int> ptr(n + 1, 0), col, brow, diag(n, -1), tptr{0}, tcode(nt);
  std::vector<double> tw(nt);
  int64_t t_out = 0;
  for (int i = 0; i < n; i++) {
    auto b0 = terms.begin() + rcnt[i], b1 = terms.begin() + rcnt[i + 1];
    std::sort(b0, b1, [](const Term &x, const Term &y) {
      return x.j != y.j ? x.j < y.j : x.code < y.code;
    });
    for (auto it = b0; it != b1; ++it) {
      if (it == b0 || it->j != (it - 1)->j) {
        if (it != b0) tptr.push_back((int)t_out);
        if (it->j == i) diag[i] = (int)col.size();
        col.push_back(it->j);
        brow.push_back(i);
      }
      tcode[t_out] = it->code;
      tw[t_out++] = it->w;
    }
    if (b1 != b0) tptr.push_back((int)t_out);
    ptr[i + 1] = (int)col.size();
    if (diag[i] < 0) {
      last_error = "galerkin: node " + std::to_string(i) + " of level " + std::to_string(l) +
                   " touches no element";
      return -1;
    }
  }
  L.nnzb = (int)col.size();
  L.bsr_ptr_h = ptr;
  L.bsr_col_h = col;
  if (L.bsr_ptr.upload(ptr) || L.bsr_col.upload(col) || L.bsr_row.upload(brow) ||
      L.bsr_diag.upload(diag) || L.bt_ptr.upload(tptr) || L.bt_code.upload(tcode) ||
      L.bt_w.upload(tw) || L.bsr64.alloc(9 * (size_t)L.nnzb) || L.bsr32.alloc(9 * (size_t)L.nnzb)) {
    last_error = "galerkin: device allocation failed";
    return -1;
  }
  L.has_bsr = true;
  return 0;
}

int Problem::set_galerkin(bool on) {
  if (on && !galerkin_built_ && build_galerkin()) return -1;
  if (on != galerkin_) {
    galerkin_ = on;
    have_setup_ = false;
    drop_graphs();
    gen_++;
  }
  return 0;
}

// per design: K_E level by level, then the BSR values (FP64, and FP32 for
// the mixed-precision V-cycle)
int Problem::galerkin_setup(int precision) {
  const int nl = (int)levels_.size();
  const bool nodal = design_mode_ == 1;
  for (int l = 1; l < nl; l++) {
    Level &F = levels_[l - 1], &C = levels_[l];
    const int mode = l > 1 ? 1 : (nodal ? 2 : 0);
    const double *s = l > 1 ? nullptr : (nodal ? F.sg64.p : F.s64.p);
    const double *Ktab = nodal ? Kgd_.p : K0d_.p;
    launch_galerkin_elem(mode, C.n_elem, F.sc_ptr.p, F.sc_child.p, F.g_place.p, F.d_cls.p, s, Ktab,
                         l > 1 ? F.KE.p : nullptr, C.KE.p, st_);
    if (C.has_bsr) {
      launch_bsr_assemble(bsr<double>(l), C.bt_ptr.p, C.bt_code.p, C.bt_w.p, C.bsr_row.p, C.KE.p,
                          C.bsr64.p, st_);
      if (precision == 1)
        launch_convert<double, float>(9 * (int64_t)C.nnzb, C.bsr64.p, C.bsr32.p, st_);
    }
  }
  GCHK(hipGetLastError());
  return 0;
}

int Problem::level_matrix(int l, std::vector<int> &ptr, std::vector<int> &col,
                          std::vector<double> &val) {
  const int nl = (int)levels_.size();
  // smoothed Galerkin levels, and the coarsest level (factored from its BSR
  // matrix in either mode)
  if (l < 1 || l >= nl || !levels_[l].has_bsr || (!galerkin_ && l != nl - 1) ||
      (!galerkin_ && !coarse_direct_)) {
    last_error = "level_matrix: no assembled operator on that level";
    return -1;
  }
  if (!have_setup_) {
    last_error = "level_matrix: run setup first";
    return -1;
  }
  Level &L = levels_[l];
  ptr.resize(L.n_nodes + 1);
  col.resize(L.nnzb);
  val.resize(9 * (size_t)L.nnzb);
  GCHK(hipStreamSynchronize(st_));
  GCHK(hipMemcpy(ptr.data(), L.bsr_ptr.p, ptr.size() * sizeof(int), hipMemcpyDeviceToHost));
  GCHK(hipMemcpy(col.data(), L.bsr_col.p, col.size() * sizeof(int), hipMemcpyDeviceToHost));
  GCHK(hipMemcpy(val.data(), L.bsr64.p, val.size() * sizeof(double), hipMemcpyDeviceToHost));
  return 0;
}

}  // namespace tmrx

namespace tmrx {

// Greedy colouring of level l's node graph (the BSR pattern, symmetric):
// nodes in lexicographic order of their points (z, then y, then x; on a
// uniform grid this reaches the 8 parity colours, where Morton order needs
// 10) take the smallest colour none of their coloured neighbours holds.
// Nodes of one colour share no block, so a colour's SOR updates are
// independent; the colouring is checked before it is uploaded.
int Problem::build_sor(int l) {
  Level &L = levels_[l];
  if (!L.has_bsr && build_bsr(l)) return -1;
  const int n = L.n_nodes;
  const std::vector<int> &ptr = L.bsr_ptr_h, &col = L.bsr_col_h;
  std::vector<int> visit(n);
  for (int i = 0; i < n; i++) visit[i] = i;
  const std::vector<double> &X = L.forest->points();
  const size_t nd = (size_t)L.forest->num_dep_nodes();
  if (X.size() >= 3 * (nd + (size_t)n)) {
    auto pt = [&](int i) { return &X[3 * (nd + (size_t)i)]; };
    std::stable_sort(visit.begin(), visit.end(), [&](int a, int b) {
      const double *pa = pt(a), *pb = pt(b);
      if (pa[2] != pb[2]) return pa[2] < pb[2];
      if (pa[1] != pb[1]) return pa[1] < pb[1];
      return pa[0] < pb[0];
    });
  }
  std::vector<int> colour(n, -1);
  int nc = 0;
  for (int i : visit) {
    uint64_t used = 0;
    for (int q = ptr[i]; q < ptr[i + 1]; q++) {
      const int j = col[q];
      if (j != i && colour[j] >= 0) used |= uint64_t(1) << colour[j];
    }
    if (used == ~uint64_t(0)) {
      last_error = "sor: level " + std::to_string(l) + " needs more than 64 colours";
      return -1;
    }
    int c = 0;
    while (used & (uint64_t(1) << c)) c++;
    colour[i] = c;
    nc = std::max(nc, c + 1);
  }
  for (int i = 0; i < n; i++)
    for (int q = ptr[i]; q < ptr[i + 1]; q++)
      if (col[q] != i && colour[col[q]] == colour[i]) {
        last_error = "sor: colouring of level " + std::to_string(l) + " puts coupled nodes " +
                     std::to_string(i) + " and " + std::to_string(col[q]) + " in one colour";
        return -1;
      }
  std::vector<int> cptr(nc + 1, 0), nodes(n);
  for (int i = 0; i < n; i++) cptr[colour[i] + 1]++;
  for (int c = 0; c < nc; c++) cptr[c + 1] += cptr[c];
  std::vector<int> fill(cptr.begin(), cptr.end() - 1);
  for (int i = 0; i < n; i++) nodes[fill[colour[i]]++] = i;
  // the matrix in sweep order: a colour's rows contiguous
  std::vector<int> pptr(n + 1, 0), pcol, pmap;
  pcol.reserve(col.size());
  pmap.reserve(col.size());
  for (int k = 0; k < n; k++) {
    const int i = nodes[k];
    for (int q = ptr[i]; q < ptr[i + 1]; q++) {
      pcol.push_back(col[q]);
      pmap.push_back(q);
    }
    pptr[k + 1] = (int)pcol.size();
  }
  if (L.sor_nodes.upload(nodes) || L.sor_ptr.upload(pptr) || L.sor_col.upload(pcol) ||
      L.sor_map.upload(pmap) || L.sor_val64.alloc(9 * (size_t)L.nnzb) ||
      L.sor_val32.alloc(9 * (size_t)L.nnzb)) {
    last_error = "sor: device allocation failed";
    return -1;
  }
  L.sor_cptr = cptr;
  L.sor_nodes_h = nodes;
  L.has_sor = true;
  return 0;
}

// The matrix SOR smooths at level l, assembled at every setup: a Galerkin
// level's P^T A P (galerkin_setup has formed it), else the scaled element
// matrices of the level scattered into its BSR pattern; FP32 values for
// the mixed-precision V-cycle.
int Problem::sor_assemble(int l, int precision) {
  Level &L = levels_[l];
  if (!L.has_sor && build_sor(l)) return -1;
  if (!(galerkin_ && l > 0)) {
    const bool nodal = design_mode_ == 1;
    if ((L.KE.n < 576 * (size_t)L.n_elem && L.KE.alloc(576 * (size_t)L.n_elem)) ||
        (nodal && Kgd_.n == 0 && Kgd_.upload(Kgh_))) {
      last_error = "sor: device allocation failed";
      return -1;
    }
    launch_elem_matrices(L.n_elem, L.d_cls.p, nodal ? L.sg64.p : L.s64.p, nodal ? Kgd_.p : K0d_.p,
                         nodal, L.KE.p, st_);
    launch_bsr_assemble(bsr<double>(l), L.bt_ptr.p, L.bt_code.p, L.bt_w.p, L.bsr_row.p, L.KE.p,
                        L.bsr64.p, st_);
  }
  // (Galerkin levels: galerkin_setup has filled bsr64) -> sweep order
  launch_bsr_gather(L.nnzb, L.sor_map.p, L.bsr64.p, L.sor_val64.p, st_);
  if (precision == 1)
    launch_convert<double, float>(9 * (int64_t)L.nnzb, L.sor_val64.p, L.sor_val32.p, st_);
  GCHK(hipGetLastError());
  return 0;
}

template <class T>
void Problem::sor_smooth(int l, const T *b, T *x, int nrhs, bool from_zero) {
  Level &L = levels_[l];
  if (from_zero) launch_fill<T>(L.ndof * (int64_t)nrhs, T(0), x, st_);
  BSR<T> A = bsr<T>(l);
  A.ptr = L.sor_ptr.p;
  A.col = L.sor_col.p;
  if constexpr (sizeof(T) == 8)
    A.val = L.sor_val64.p;
  else
    A.val = L.sor_val32.p;
  A.run = run_;
  const int nc = (int)L.sor_cptr.size() - 1;
  launch_sor_sweep<T>(A, L.sor_nodes.p, L.sor_cptr.data(), nc, false, b, x, omega_, st_, nrhs,
                      L.ndof);
  launch_sor_sweep<T>(A, L.sor_nodes.p, L.sor_cptr.data(), nc, true, b, x, omega_, st_, nrhs,
                      L.ndof);
}

template void Problem::sor_smooth<float>(int, const float *, float *, int, bool);
template void Problem::sor_smooth<double>(int, const double *, double *, int, bool);

int Problem::set_smoother(bool sor, double omega) {
  if (!(omega > 0.0 && omega < 2.0)) {
    last_error = "mg: omega must lie in (0, 2)";
    return -1;
  }
  if (sor != sor_ || omega != omega_) {
    sor_ = sor;
    omega_ = omega;
    have_setup_ = false;
    drop_graphs();
    gen_++;
  }
  return 0;
}

int Problem::smoother_order(int l, std::vector<int> &nodes, std::vector<int> &cptr) {
  if (l < 0 || l >= (int)levels_.size() || !sor_ || !levels_[l].has_sor) {
    last_error = "smoother_order: no SOR ordering on that level (set SOR smoothing and run a setup)";
    return -1;
  }
  nodes = levels_[l].sor_nodes_h;
  cptr = levels_[l].sor_cptr;
  return 0;
}

}  // namespace tmrx
