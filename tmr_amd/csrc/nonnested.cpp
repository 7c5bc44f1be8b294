// Coarse spaces that are not nested: interpolation and element ownership
// between two forests of box blocks with unrelated block lattices over the
// same region (a cube of 13 elements per edge, say, under one of 7).  The
// reference builds its multigrid hierarchies with coarsen() only
// (TopOptUtils.createTopoProblem, tmr/TopOptUtils.py:69-99), which keeps
// level-0 octants (src/TMROctForest.cpp:2119-2164): a mesh whose edges have
// no power-of-two factor has no coarse level there.  These routines give
// such meshes one by locating points physically instead of by octant
// containment.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <string>
#include <utility>
#include <vector>

#include "forest.h"

namespace tmrx {

namespace {

// Uniform grid over the blocks' bounding boxes: the lowest-index block that
// contains a point, with a tolerance of 1e-10 of the block's size so that
// points on shared faces are found from either side.
struct BlockGrid {
  double g0[3] = {0, 0, 0}, cs = 1.0;
  int64_t nc[3] = {1, 1, 1};
  std::vector<double> lo, hi, tol;
  std::vector<int> ptr, idx;

  int64_t cell(int a, double v) const {
    const double c = std::floor((v - g0[a]) / cs);
    if (!(c >= 0.0)) return 0;
    return std::min<int64_t>((int64_t)c, nc[a] - 1);
  }

  void build(const BlockConn &bd, const std::vector<double> &Xv) {
    const int nb = bd.num_blocks;
    lo.assign(3 * (size_t)nb, 0.0);
    hi.assign(3 * (size_t)nb, 0.0);
    tol.assign(nb, 0.0);
    double G0[3] = {HUGE_VAL, HUGE_VAL, HUGE_VAL}, G1[3] = {-HUGE_VAL, -HUGE_VAL, -HUGE_VAL};
    double vol = 0.0;
    for (int b = 0; b < nb; b++) {
      const double *p0 = &Xv[3 * (size_t)bd.block_conn[8 * (size_t)b]];
      const double *p7 = &Xv[3 * (size_t)bd.block_conn[8 * (size_t)b + 7]];
      double ext = 0.0, v = 1.0;
      for (int a = 0; a < 3; a++) {
        lo[3 * (size_t)b + a] = p0[a];
        hi[3 * (size_t)b + a] = p7[a];
        G0[a] = std::min(G0[a], p0[a]);
        G1[a] = std::max(G1[a], p7[a]);
        ext = std::max(ext, p7[a] - p0[a]);
        v *= p7[a] - p0[a];
      }
      tol[b] = 1e-10 * ext;
      vol += v;
    }
    if (nb == 0) return;
    for (int a = 0; a < 3; a++) g0[a] = G0[a];
    // cells about the size of a mean block, at most 8 per block
    cs = std::cbrt(vol / nb);
    for (;;) {
      double tot = 1.0;
      for (int a = 0; a < 3; a++) {
        nc[a] = std::max<int64_t>(1, (int64_t)std::ceil((G1[a] - G0[a]) / cs));
        tot *= (double)nc[a];
      }
      if (tot <= 8.0 * nb + 64.0) break;
      cs *= 1.5;
    }
    const size_t ncell = (size_t)(nc[0] * nc[1] * nc[2]);
    std::vector<int> cnt(ncell + 1, 0);
    auto each_cell = [&](int b, auto fn) {
      int64_t c0[3], c1[3];
      for (int a = 0; a < 3; a++) {
        c0[a] = cell(a, lo[3 * (size_t)b + a] - tol[b]);
        c1[a] = cell(a, hi[3 * (size_t)b + a] + tol[b]);
      }
      for (int64_t k = c0[2]; k <= c1[2]; k++)
        for (int64_t j = c0[1]; j <= c1[1]; j++)
          for (int64_t i = c0[0]; i <= c1[0]; i++) fn((size_t)(i + nc[0] * (j + nc[1] * k)));
    };
    for (int b = 0; b < nb; b++) each_cell(b, [&](size_t c) { cnt[c + 1]++; });
    for (size_t c = 0; c < ncell; c++) cnt[c + 1] += cnt[c];
    ptr = cnt;
    idx.assign((size_t)ptr[ncell], 0);
    std::vector<int> fill(ptr.begin(), ptr.end() - 1);
    for (int b = 0; b < nb; b++) each_cell(b, [&](size_t c) { idx[fill[c]++] = b; });
  }

  // lowest-index block containing p (ascending within each cell), or -1
  int find(const double p[3]) const {
    if (ptr.empty()) return -1;
    const size_t c = (size_t)(cell(0, p[0]) + nc[0] * (cell(1, p[1]) + nc[1] * cell(2, p[2])));
    for (int j = ptr[c]; j < ptr[c + 1]; j++) {
      const int b = idx[j];
      bool in = true;
      for (int a = 0; a < 3; a++)
        if (p[a] < lo[3 * (size_t)b + a] - tol[b] || p[a] > hi[3 * (size_t)b + a] + tol[b]) in = false;
      if (in) return b;
    }
    return -1;
  }
};

inline double clamp01(double v) { return v < 0.0 ? 0.0 : (v > 1.0 ? 1.0 : v); }

}  // namespace

bool Forest::same_blocks(const Forest *o) const {
  if (!o || bd_.num_blocks != o->bd_.num_blocks || bd_.block_conn != o->bd_.block_conn ||
      Xv_.size() != o->Xv_.size())
    return false;
  for (size_t i = 0; i < Xv_.size(); i++)
    if (std::fabs(Xv_[i] - o->Xv_[i]) > 1e-12 * (1.0 + std::fabs(Xv_[i]))) return false;
  return true;
}

// The coarse element containing physical point p: block b, octant index
// (returned), and p's parameter u in [0, 1]^3 within the block.
static int locate_point(const Forest &coarse, const BlockGrid &grid, const double p[3],
                        double u[3]) {
  const int b = grid.find(p);
  if (b < 0) return -1;
  int32_t ic[3];
  for (int a = 0; a < 3; a++) {
    const double l = grid.lo[3 * (size_t)b + a], h = grid.hi[3 * (size_t)b + a];
    u[a] = clamp01((p[a] - l) / (h - l));
    const double s = std::floor(u[a] * (double)HMAX);
    ic[a] = (int32_t)std::min<double>(std::max(s, 0.0), (double)(HMAX - 1));
  }
  const Oct q{b, ic[0], ic[1], ic[2], 0, (int16_t)MAX_LEVEL, 0};
  return coarse.locate(q);
}

int Forest::create_interpolation_geometric(Forest *coarse, InterpCSR &P) {
  if (!coarse || coarse == this) {
    last_error = "create_interpolation_geometric: need a different coarse forest";
    return -1;
  }
  if (!blocks_are_boxes() || !coarse->blocks_are_boxes()) {
    last_error = "create_interpolation_geometric: both forests need axis-aligned box blocks with points";
    return -1;
  }
  if (create_nodes() != 0) return -1;
  if (coarse->create_nodes() != 0) {
    last_error = "create_interpolation_geometric: coarse forest: " + coarse->last_error;
    return -1;
  }
  BlockGrid grid;
  grid.build(coarse->bd_, coarse->Xv_);
  const auto &co = coarse->octs_;
  const auto &cc = coarse->conn_;
  const auto &cdp = coarse->dep_ptr_;
  const auto &cdc = coarse->dep_conn_;
  const auto &cdw = coarse->dep_w_;
  P.nrows = num_owned_;
  P.ncols = coarse->num_owned_;
  P.rowp.assign((size_t)P.nrows + 1, 0);
  P.cols.clear();
  P.w.clear();
  std::vector<std::pair<int, double>> wts;
  for (int i = 0; i < num_owned_; i++) {
    const double *p = &X_[3 * ((size_t)i + num_dep_)];
    double u[3];
    const int E = locate_point(*coarse, grid, p, u);
    if (E < 0) {
      last_error = "create_interpolation_geometric: fine node " + std::to_string(i) +
                   " lies outside the coarse forest";
      return -1;
    }
    const Oct &o = co[E];
    const double hE = (double)(1 << (MAX_LEVEL - o.level));
    const double t[3] = {clamp01((u[0] * HMAX - o.x) / hE), clamp01((u[1] * HMAX - o.y) / hE),
                         clamp01((u[2] * HMAX - o.z) / hE)};
    wts.clear();
    for (int k = 0; k < 8; k++) {
      double N = 1.0;
      for (int a = 0; a < 3; a++) N *= ((k >> a) & 1) ? t[a] : 1.0 - t[a];
      if (N == 0.0) continue;
      const int g = cc[8 * (size_t)E + k];
      if (g >= 0) {
        wts.push_back({g, N});
      } else {
        for (int j = cdp[-g - 1]; j < cdp[-g]; j++) wts.push_back({cdc[j], N * cdw[j]});
      }
    }
    std::sort(wts.begin(), wts.end(),
              [](const std::pair<int, double> &a, const std::pair<int, double> &b) { return a.first < b.first; });
    const size_t start = P.cols.size();
    for (auto &q : wts) {
      if (P.cols.size() > start && P.cols.back() == q.first)
        P.w.back() += q.second;
      else {
        P.cols.push_back(q.first);
        P.w.push_back(q.second);
      }
    }
    // rounding leftovers of corners the point lies on the far side of
    size_t keep = start;
    for (size_t j = start; j < P.cols.size(); j++)
      if (std::fabs(P.w[j]) > 1e-13) {
        P.cols[keep] = P.cols[j];
        P.w[keep++] = P.w[j];
      }
    P.cols.resize(keep);
    P.w.resize(keep);
    P.rowp[(size_t)i + 1] = (int)P.cols.size();
  }
  return 0;
}

int Forest::locate_centres(Forest *coarse, std::vector<int> &owner) {
  if (!coarse || !blocks_are_boxes() || !coarse->blocks_are_boxes()) {
    last_error = "locate_centres: both forests need axis-aligned box blocks with points";
    return -1;
  }
  BlockGrid grid;
  grid.build(coarse->bd_, coarse->Xv_);
  owner.assign(octs_.size(), -1);
  for (size_t e = 0; e < octs_.size(); e++) {
    const Oct &o = octs_[e];
    const double h = (double)(1 << (MAX_LEVEL - o.level));
    const double *p0 = &Xv_[3 * (size_t)bd_.block_conn[8 * (size_t)o.block]];
    const double *p7 = &Xv_[3 * (size_t)bd_.block_conn[8 * (size_t)o.block + 7]];
    const double s[3] = {(o.x + 0.5 * h) / HMAX, (o.y + 0.5 * h) / HMAX, (o.z + 0.5 * h) / HMAX};
    double p[3], u[3];
    for (int a = 0; a < 3; a++) p[a] = p0[a] + s[a] * (p7[a] - p0[a]);
    owner[e] = locate_point(*coarse, grid, p, u);
  }
  return 0;
}

}  // namespace tmrx
