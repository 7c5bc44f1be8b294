// Multigrid-preconditioned FP64 flexible PCG on one GPU.  See problem.h.
#include "problem.h"

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstring>
#include <functional>
#include <limits>
#include <numeric>

#include "fe.h"

namespace tmrx {

#define HIPCHK(x)                                                       \
  do {                                                                  \
    hipError_t e_ = (x);                                                \
    if (e_ == hipSuccess) e_ = take_launch_error();                     \
    if (e_ != hipSuccess) {                                             \
      last_error = std::string("HIP error: ") + hipGetErrorString(e_) + \
                   " at " + #x;                                         \
      return -1;                                                        \
    }                                                                   \
  } while (0)

// Tile size of a coarse level: the finest level's size, reduced (down to 64
// nodes) until the level has at least 2048 tiles, so that small levels still
// spread over the chip and each tile's serial latency stays short.
static int coarse_tile_nodes(int fine_tile, int n_nodes) {
  int t = fine_tile;
  while (t > 64 && (int64_t)n_nodes < 2048LL * t) t /= 2;
  return t;
}

void Problem::drop_graphs() {
  for (int k = 0; k < 2; k++)
    if (gexec_[k]) {
      (void)hipGraphExecDestroy(gexec_[k]);
      gexec_[k] = nullptr;
    }
  g_gen_ = -1;
}

Problem::~Problem() {
  drop_graphs();
  dense_blas_destroy(blas_);
  for (int k = 0; k < 2; k++)
    if (gev_[k]) (void)hipEventDestroy(gev_[k]);
  if (cap_st_) (void)hipStreamDestroy(cap_st_);
  if (nact_host_) (void)hipHostFree(nact_host_);
  for (auto &e : ev_pool_) {
    (void)hipEventDestroy(e.first);
    (void)hipEventDestroy(e.second);
  }
}

int Problem::kernel_timing(int enable) {
  timing_ = enable;
  ev_used_ = 0;
  ev_prec_mask_ = 0;
  return 0;
}

int Problem::kernel_timing_result(double *total_ms, int *count, int *mask) {
  HIPCHK(hipStreamSynchronize(st_));
  double t = 0.0;
  for (size_t i = 0; i < ev_used_; i++) {
    float ms = 0.0f;
    HIPCHK(hipEventElapsedTime(&ms, ev_pool_[i].first, ev_pool_[i].second));
    t += ms;
  }
  if (total_ms) *total_ms = t;
  if (count) *count = (int)ev_used_;
  if (mask) *mask = ev_prec_mask_;
  return 0;
}

template <>
LevelVecs<float> &Problem::vecs<float>(int l) { return levels_[l].v32; }
template <>
LevelVecs<double> &Problem::vecs<double>(int l) { return levels_[l].v64; }
template <>
const float *Problem::K0<float>() { return K0f_.p; }
template <>
const double *Problem::K0<double>() { return K0d_.p; }
template <>
const float *Problem::scale<float>(int l) { return levels_[l].s32.p; }
template <>
const double *Problem::scale<double>(int l) { return levels_[l].s64.p; }
template <>
const float *Problem::rtab<float>() { return R32_.p; }
template <>
const double *Problem::rtab<double>() { return R64_.p; }
template <>
const float *Problem::gtab<float>() { return G32_.p; }
template <>
const double *Problem::gtab<double>() { return G64_.p; }
template <>
const float *Problem::gscale<float>(int l) { return levels_[l].sg32.p; }
template <>
const double *Problem::gscale<double>(int l) { return levels_[l].sg64.p; }

// ---------------------------------------------------------------------------
int Problem::init(int nlevels, Forest **forests, const Material &mat, int nbcs,
                  const char **names, const int *dof_masks) {
  if (nlevels < 1) {
    last_error = "problem: need at least one level";
    return -1;
  }
  mat_ = mat;
  levels_.clear();
  levels_.resize(nlevels);
  for (int l = 0; l < nlevels; l++) {
    Forest *f = forests[l];
    if (!f) {
      last_error = "problem: null forest";
      return -1;
    }
    if (!f->blocks_are_boxes()) {
      last_error = "problem: every block must be an axis-aligned box with points set";
      return -1;
    }
    // nodes and (below) interpolations on the problem's GPU: the same
    // results as the host passes, bit for bit
    int dev = 0;
    (void)hipGetDevice(&dev);
    if (f->create_nodes_device(dev)) {
      last_error = "problem: level " + std::to_string(l) + ": " + f->last_error;
      return -1;
    }
    levels_[l].forest = f;
  }
  if (find_components()) return -1;
  // stiffness classes: blocks with the same box extents share K0 (over
  // every level: the levels of a hierarchy that is not nested have blocks
  // of their own)
  std::vector<std::array<double, 3>> ext;
  std::vector<std::vector<int>> block_cls(nlevels);
  for (int l = 0; l < nlevels; l++) {
    Forest *fl = forests[l];
    const BlockConn &bd = fl->conn_data();
    if (l > 0 && fl->same_blocks(forests[l - 1])) {
      block_cls[l] = block_cls[l - 1];
      continue;
    }
    block_cls[l].assign(bd.num_blocks, 0);
    for (int b = 0; b < bd.num_blocks; b++) {
      Oct root{b, 0, 0, 0, 0, 0, 0};
      double h[3];
      fl->element_extent(root, h);
      int c = -1;
      for (size_t k = 0; k < ext.size(); k++) {
        bool eq = true;
        for (int a = 0; a < 3; a++)
          if (std::fabs(ext[k][a] - h[a]) > 1e-12 * std::fabs(h[a])) eq = false;
        if (eq) {
          c = (int)k;
          break;
        }
      }
      if (c < 0) {
        c = (int)ext.size();
        ext.push_back({h[0], h[1], h[2]});
      }
      block_cls[l][b] = c;
    }
  }
  nclasses_ = (int)ext.size();
  if (nclasses_ > 256) {  // element records carry the class in one byte
    last_error = "problem: more than 256 distinct block extents";
    return -1;
  }
  double C[36];
  isotropic_C(mat.E, mat.nu, C);
  K0h_.assign(576 * (size_t)nclasses_, 0.0);
  for (int c = 0; c < nclasses_; c++)
    hex8_box_stiffness(ext[c][0], ext[c][1], ext[c][2], C, &K0h_[576 * c]);
  // K0 in the reflection-symmetric basis: with H the 8-point Walsh-Hadamard
  // matrix over the corners, (H K0_cc' H)[t][t'] vanishes unless
  // t ^ (1 << c) == t' ^ (1 << c'); the surviving entries form one 3x3 block
  // per s = t ^ (1 << c).  Checked here: a box element with an isotropic
  // material is symmetric under the three coordinate reflections.
  {
    std::vector<double> rt(REFL_TABLE * (size_t)nclasses_, 0.0);
    auto Hs = [](int t, int a) { return (__builtin_popcount(t & a) & 1) ? -1.0 : 1.0; };
    for (int cl = 0; cl < nclasses_; cl++) {
      const double *K = &K0h_[576 * (size_t)cl];
      double kmax = 0.0, off = 0.0;
      for (int i = 0; i < 576; i++) kmax = std::max(kmax, std::fabs(K[i]));
      for (int c = 0; c < 3; c++)
        for (int cp = 0; cp < 3; cp++)
          for (int t = 0; t < 8; t++)
            for (int tp = 0; tp < 8; tp++) {
              double v = 0.0;
              for (int a = 0; a < 8; a++)
                for (int b = 0; b < 8; b++) v += Hs(t, a) * K[24 * (3 * a + c) + 3 * b + cp] * Hs(tp, b);
              const int sv = t ^ (1 << c);
              if (sv == (tp ^ (1 << cp)) && refl_nz(sv, 3 * c + cp))
                rt[REFL_TABLE * (size_t)cl + 9 * sv + 3 * c + cp] = v / 64.0;
              else  // off the blocks, or a structural zero the kernels skip
                off = std::max(off, std::fabs(v));
            }
      if (off > 1e-12 * 64.0 * kmax) {
        last_error = "problem: element stiffness is not symmetric under the coordinate reflections";
        return -1;
      }
    }
    std::vector<float> rtf(rt.begin(), rt.end());
    if (R64_.upload(rt) || R32_.upload(rtf)) {
      last_error = "problem: device allocation failed";
      return -1;
    }
  }
  // nodal design: Gauss-point stiffness (host, coarse factor) and gradient
  // tables with lambda |J|, mu |J| (device)
  {
    Kgh_.assign(8 * 576 * (size_t)nclasses_, 0.0);
    std::vector<double> gt(NODAL_TABLE * (size_t)nclasses_, 0.0);
    const double lam = mat.E * mat.nu / ((1.0 + mat.nu) * (1.0 - 2.0 * mat.nu));
    const double mu = mat.E / (2.0 * (1.0 + mat.nu));
    for (int c = 0; c < nclasses_; c++) {
      for (int q = 0; q < 8; q++)
        hex8_box_stiffness_gp(ext[c][0], ext[c][1], ext[c][2], C, q, &Kgh_[576 * (8 * (size_t)c + q)]);
      double *g = &gt[NODAL_TABLE * (size_t)c];
      hex8_box_grad_table(ext[c][0], ext[c][1], ext[c][2], g);
      const double wj = ext[c][0] * ext[c][1] * ext[c][2] / 8.0;
      g[192] = lam * wj;
      g[193] = mu * wj;
    }
    std::vector<float> gtf(gt.begin(), gt.end());
    if (G64_.upload(gt) || G32_.upload(gtf)) {
      last_error = "problem: device allocation failed";
      return -1;
    }
  }
  // Gauss-point stress operators C0 B_g (6 x 24, 8 points) and volumes of the
  // full-block element of every class (p-norm stress functional)
  {
    std::vector<double> cb(8 * 144 * (size_t)nclasses_), vol(nclasses_);
    const double gp = 1.0 / std::sqrt(3.0);
    for (int c = 0; c < nclasses_; c++) {
      vol[c] = ext[c][0] * ext[c][1] * ext[c][2];
      for (int g = 0; g < 8; g++) {
        const double pt[3] = {(g & 1) ? gp : -gp, (g & 2) ? gp : -gp, (g & 4) ? gp : -gp};
        double B[144], N[8];
        hex8_box_B(ext[c][0], ext[c][1], ext[c][2], pt, B, N);
        for (int i = 0; i < 6; i++)
          for (int j = 0; j < 24; j++) {
            double t = 0.0;
            for (int k = 0; k < 6; k++) t += C[6 * i + k] * B[24 * k + j];
            cb[144 * (8 * (size_t)c + g) + 24 * i + j] = t;
          }
      }
    }
    if (CB_.upload(cb) || cvol_.upload(vol)) {
      last_error = "problem: device allocation failed";
      return -1;
    }
  }
  std::vector<float> K0f(K0h_.begin(), K0h_.end());
  // MFMA A fragments of the FP32 K0 for the matrix-core tile apply:
  // [class][part hi, mid, lo][row block][lane][8 bf16].  Lane l holds row
  // l % 16 of the block and K columns 8 (l / 16) .. +7.  Row 4q + j of block
  // b is component j of corner 2q + b; column 8q + 4c + j is component j of
  // corner 2q + c; j = 3 is zero padding.
  std::vector<uint16_t> af(3072 * (size_t)nclasses_, 0);
  auto bf16 = [](float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
  };
  auto bf16f = [](uint16_t h) {
    const uint32_t u = (uint32_t)h << 16;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
  };
  for (int c = 0; c < nclasses_; c++)
    for (int blk = 0; blk < 2; blk++)
      for (int l = 0; l < 64; l++)
        for (int v = 0; v < 8; v++) {
          const int i = l % 16, q = i / 4, j = i % 4;
          const int gk = l / 16, cc = v / 4, jj = v % 4;
          float k = 0.0f;
          if (j < 3 && jj < 3) {
            const int row = 3 * (2 * q + blk) + j, col = 3 * (2 * gk + cc) + jj;
            k = K0f[576 * c + 24 * row + col];
          }
          const uint16_t hi = bf16(k);
          const float r = k - bf16f(hi);
          const uint16_t mid = bf16(r);
          const uint16_t lo = bf16(r - bf16f(mid));
          const size_t base = 3072 * (size_t)c + 8 * (size_t)l + v;
          af[base + 512 * (0 + blk)] = hi;
          af[base + 512 * (2 + blk)] = mid;
          af[base + 512 * (4 + blk)] = lo;
        }
  // FP32 A fragments for the FP32-input matrix-core form: lane l of row
  // block b, step s holds K0[row][dof] with row 16 b + l % 16 = 4 a + j
  // (component j of corner a; j = 3 pads with zero) and dof 4 s + l / 16
  std::vector<float> af32(768 * (size_t)nclasses_, 0.0f);
  for (int c = 0; c < nclasses_; c++)
    for (int blk = 0; blk < 2; blk++)
      for (int st = 0; st < 6; st++)
        for (int l = 0; l < 64; l++) {
          const int row = 16 * blk + l % 16, a = row / 4, j = row % 4, dof = 4 * st + l / 16;
          af32[768 * (size_t)c + (blk * 6 + st) * 64 + l] =
              j < 3 ? K0f[576 * (size_t)c + 24 * (3 * a + j) + dof] : 0.0f;
        }
  if (K0d_.upload(K0h_) || K0f_.upload(K0f) || afrag_.upload(af) || afrag32_.upload(af32)) {
    last_error = "problem: device allocation failed";
    return -1;
  }
  for (int l = 0; l < nlevels; l++) {
    Level &L = levels_[l];
    const auto &octs = L.forest->octants();
    L.cls.resize(octs.size());
    for (size_t e = 0; e < octs.size(); e++) L.cls[e] = block_cls[l][octs[e].block];
    if (build_level(l, nbcs, names, dof_masks)) return -1;
    if (l > 0) L.tile_nodes = coarse_tile_nodes(levels_[0].tile_nodes, L.n_nodes);
    if (build_tiles(l)) return -1;
  }
  for (int l = 0; l + 1 < nlevels; l++)
    if (build_transfer(l)) return -1;
  const int64_t n = levels_[0].ndof;
  if (build_plan()) return -1;
  if (ensure_batch(1) || rho_.alloc(levels_[0].n_elem) || energy_.alloc(levels_[0].n_elem) ||
      psi_.alloc(n) || adjf_.alloc(n) || dfdrho_.alloc(levels_[0].n_elem)) {
    last_error = "problem: device allocation failed";
    return -1;
  }
  HIPCHK(hipMemset(u_.p, 0, n * sizeof(double)));
  {
    std::vector<uint8_t> z(levels_[0].n_nodes, 0);
    if (nomask_.upload(z)) {
      last_error = "problem: device allocation failed";
      return -1;
    }
  }
  return 0;
}

// Prescribed values of the clamped dofs: TMRBoundaryConditions' bc_vals
// (src/TMRBoundaryConditions.cpp:68-116) handed to TACSAssembler::addBCs by
// TMROctTACSCreator::setBoundaryConditions (src/TMR_TACSCreator.cpp:468-500).
// The solve lifts them: K_ff u_f = f_f - K_fc u_c, u_c = the values.
int Problem::set_bc_values(const double *ubc, int n, bool on_device) {
  if (levels_.empty()) {
    last_error = "problem: not initialised";
    return -1;
  }
  const Level &L = levels_[0];
  if (n != L.ndof) {
    last_error = "set_bc_values: expected one value per finest-level dof";
    return -1;
  }
  std::vector<double> h(n);
  if (on_device) {
    HIPCHK(hipStreamSynchronize(st_));
    HIPCHK(hipMemcpy(h.data(), ubc, n * sizeof(double), hipMemcpyDeviceToHost));
  } else {
    std::copy(ubc, ubc + n, h.begin());
  }
  bool any = false;
  for (int i = 0; i < L.n_nodes; i++)
    for (int c = 0; c < 3; c++) {
      double &v = h[3 * (size_t)i + c];
      if (!((L.bcmask[i] >> c) & 1)) v = 0.0;
      if (v != 0.0) any = true;
    }
  if (ubc_.upload(h)) {
    last_error = "set_bc_values: device allocation failed";
    return -1;
  }
  has_ubc_ = any;
  have_solution_ = false;
  return 0;
}

int Problem::build_level(int l, int nbcs, const char **names, const int *masks) {
  Level &L = levels_[l];
  Forest *f = L.forest;
  const auto &octs = f->octants();
  L.n_elem = f->num_elements();
  L.n_nodes = f->num_owned_nodes();
  L.n_dep = f->num_dep_nodes();
  L.ndof = 3 * (int64_t)L.n_nodes;
  L.conn = f->conn();
  L.lev.resize(L.n_elem);
  for (int e = 0; e < L.n_elem; e++) L.lev[e] = octs[e].level;
  // boundary conditions by name (TMROctTACSCreator::setBoundaryConditions,
  // src/TMR_TACSCreator.cpp:468-500); hanging nodes follow their parents
  L.bcmask.assign(L.n_nodes, 0);
  for (int k = 0; k < nbcs; k++) {
    std::vector<int> nodes;
    f->nodes_with_name(names[k], nodes);
    const uint8_t m = (uint8_t)((masks ? masks[k] : 7) & 7);
    for (int n : nodes)
      if (n >= 0) L.bcmask[n] |= m;
  }
  // element batches
  const int ne = L.n_elem;
  const int next = L.n_nodes + L.n_dep;
  std::vector<int> stamp(next, -1), local(next, 0);
  std::vector<int> wg_node_ptr{0}, wg_nodes, wg_elem_ptr{0}, wg_class;
  std::vector<uint16_t> lconn(8 * (size_t)ne);
  std::vector<uint8_t> color(ne);
  std::vector<uint32_t> nodecol;
  int max_nodes = 0, ncolors = 0;
  int e = 0, b = 0;
  while (e < ne) {
    const int start = e, c = L.cls[e];
    const int nb0 = (int)wg_nodes.size();
    while (e < ne && e - start < WG_ELEMS && L.cls[e] == c) {
      for (int a = 0; a < 8; a++) {
        const int g = L.conn[8 * e + a];
        const int x = g >= 0 ? g : L.n_nodes + (-g - 1);
        if (stamp[x] != b) {
          stamp[x] = b;
          local[x] = (int)wg_nodes.size() - nb0;
          wg_nodes.push_back(g);
        }
        lconn[8 * (size_t)e + a] = (uint16_t)local[x];
      }
      e++;
    }
    const int nn = (int)wg_nodes.size() - nb0;
    max_nodes = std::max(max_nodes, nn);
    nodecol.assign(nn, 0u);
    for (int k = start; k < e; k++) {
      uint32_t used = 0;
      for (int a = 0; a < 8; a++) used |= nodecol[lconn[8 * (size_t)k + a]];
      int col = 0;
      while (used & (1u << col)) col++;
      if (col >= MAX_COLORS) {
        last_error = "problem: element batch needs too many colors";
        return -1;
      }
      color[k] = (uint8_t)col;
      ncolors = std::max(ncolors, col + 1);
      for (int a = 0; a < 8; a++) nodecol[lconn[8 * (size_t)k + a]] |= 1u << col;
    }
    wg_node_ptr.push_back((int)wg_nodes.size());
    wg_elem_ptr.push_back(e);
    wg_class.push_back(c);
    b++;
  }
  // assembly CSR: node <- its slots (w = 1) and the slots of the hanging
  // nodes it supports (w = hanging-node weight), in increasing slot order
  const auto &dptr = f->dep_ptr();
  const auto &dconn = f->dep_conn();
  const auto &dw = f->dep_weights();
  const int nslots = (int)wg_nodes.size();
  std::vector<int> aptr(L.n_nodes + 1, 0);
  for (int s = 0; s < nslots; s++) {
    const int g = wg_nodes[s];
    if (g >= 0) {
      aptr[g + 1]++;
    } else {
      const int d = -g - 1;
      for (int j = dptr[d]; j < dptr[d + 1]; j++) aptr[dconn[j] + 1]++;
    }
  }
  for (int i = 0; i < L.n_nodes; i++) aptr[i + 1] += aptr[i];
  std::vector<int> aslot(aptr[L.n_nodes]);
  std::vector<float> aw(aptr[L.n_nodes]);
  {
    std::vector<int> fill(aptr.begin(), aptr.end() - 1);
    for (int s = 0; s < nslots; s++) {
      const int g = wg_nodes[s];
      if (g >= 0) {
        aslot[fill[g]] = s;
        aw[fill[g]++] = 1.0f;
      } else {
        const int d = -g - 1;
        for (int j = dptr[d]; j < dptr[d + 1]; j++) {
          const int p = dconn[j];
          aslot[fill[p]] = s;
          aw[fill[p]++] = (float)dw[j];
        }
      }
    }
  }
  // Diagonal of A = Z^T K Z beyond the per-corner terms w^2 K_e[a][a] that
  // the slot gather forms: pairs of distinct corners a < b of one element
  // whose constraint rows share an independent node i add
  // 2 Z_ai Z_bi K_e[a][b] (per dof direction) to A_ii.  Only elements with
  // hanging corners have such pairs.  Records grouped by node i:
  // code = 64 e + 8 a + b, weight 2 Z_ai Z_bi.
  {
    std::vector<std::array<int, 3>> rec;  // node, code, index into wt
    std::vector<double> wt;
    std::vector<std::pair<int, double>> lists[8];
    for (int e = 0; e < ne; e++) {
      bool hang = false;
      for (int a = 0; a < 8; a++) hang |= L.conn[8 * e + a] < 0;
      if (!hang) continue;
      for (int a = 0; a < 8; a++) {
        lists[a].clear();
        const int g = L.conn[8 * e + a];
        if (g >= 0) {
          lists[a].push_back({g, 1.0});
        } else {
          const int d = -g - 1;
          for (int j = dptr[d]; j < dptr[d + 1]; j++) lists[a].push_back({dconn[j], dw[j]});
        }
      }
      for (int a = 0; a < 8; a++)
        for (int b2 = a + 1; b2 < 8; b2++)
          for (const auto &pa : lists[a])
            for (const auto &pb : lists[b2])
              if (pa.first == pb.first) {
                rec.push_back({pa.first, 64 * e + 8 * a + b2, (int)wt.size()});
                wt.push_back(2.0 * pa.second * pb.second);
              }
    }
    std::stable_sort(rec.begin(), rec.end(),
                     [](const std::array<int, 3> &x, const std::array<int, 3> &y) { return x[0] < y[0]; });
    std::vector<int> xptr{0}, xnode, xcode;
    std::vector<double> xw;
    for (size_t k = 0; k < rec.size(); k++) {
      if (k == 0 || rec[k][0] != rec[k - 1][0]) {
        if (k) xptr.push_back((int)k);
        xnode.push_back(rec[k][0]);
      }
      xcode.push_back(rec[k][1]);
      xw.push_back(wt[rec[k][2]]);
    }
    xptr.push_back((int)rec.size());
    L.n_xd = (int)xnode.size();
    if (L.n_xd &&
        (L.xd_ptr.upload(xptr) || L.xd_node.upload(xnode) || L.xd_code.upload(xcode) ||
         L.xd_w.upload(xw))) {
      last_error = "problem: device allocation failed";
      return -1;
    }
  }
  std::vector<int> dp(dptr.begin(), dptr.end());
  if (dp.empty()) dp.push_back(0);
  std::vector<float> dwf(dw.begin(), dw.end());
  if (L.wg_node_ptr.upload(wg_node_ptr) || L.wg_nodes.upload(wg_nodes) ||
      L.wg_elem_ptr.upload(wg_elem_ptr) || L.wg_class.upload(wg_class) ||
      L.lconn.upload(lconn) || L.color.upload(color) || L.d_bcmask.upload(L.bcmask) ||
      L.dep_ptr.upload(dp) || L.dep_conn.upload(dconn) || L.dep_w.upload(dwf) ||
      L.asm_ptr.upload(aptr) || L.asm_slot.upload(aslot) || L.asm_w.upload(aw) ||
      L.d_conn.upload(L.conn) || L.d_cls.upload(L.cls) || L.d_lev.upload(L.lev) ||
      L.s64.alloc(ne) || L.s32.alloc(ne)) {
    last_error = "problem: device allocation failed";
    return -1;
  }
  ElemBatches &eb = L.eb;
  eb.n_wg = b;
  eb.n_elem = ne;
  eb.ncolors = ncolors;
  eb.max_wg_nodes = max_nodes;
  eb.wg_node_ptr = L.wg_node_ptr.p;
  eb.wg_nodes = L.wg_nodes.p;
  eb.wg_elem_ptr = L.wg_elem_ptr.p;
  eb.wg_class = L.wg_class.p;
  eb.lconn = L.lconn.p;
  eb.color = L.color.p;
  eb.dep_ptr = L.dep_ptr.p;
  eb.dep_conn = L.dep_conn.p;
  eb.dep_w = L.dep_w.p;
  eb.bcmask = L.d_bcmask.p;
  L.asmc.n_nodes = L.n_nodes;
  L.asmc.ptr = L.asm_ptr.p;
  L.asmc.slot = L.asm_slot.p;
  L.asmc.w = L.asm_w.p;
  return 0;
}

// Owner tiles: contiguous ranges of L.tile_nodes independent nodes; every
// element touching an owned node, or a hanging node with an owned parent, is
// staged with its corners in LDS slots [owned | accumulated hanging | halo].
int Problem::build_tiles(int l) {
  Level &L = levels_[l];
  Forest *f = L.forest;
  const int n = L.n_nodes, nd = L.n_dep, ne = L.n_elem;
  const auto &dptr = f->dep_ptr();
  const auto &dconn = f->dep_conn();
  const auto &dw = f->dep_weights();
  // incidence: ext node (g >= 0 -> g, dep d -> n + d) -> elements
  std::vector<int> iptr(n + nd + 1, 0), iel;
  auto xid = [&](int g) { return g >= 0 ? g : n + (-g - 1); };
  for (int e = 0; e < ne; e++)
    for (int a = 0; a < 8; a++) iptr[xid(L.conn[8 * e + a]) + 1]++;
  for (int i = 0; i < n + nd; i++) iptr[i + 1] += iptr[i];
  iel.resize(iptr[n + nd]);
  {
    std::vector<int> fill(iptr.begin(), iptr.end() - 1);
    for (int e = 0; e < ne; e++)
      for (int a = 0; a < 8; a++) iel[fill[xid(L.conn[8 * e + a])]++] = e;
  }
  // parent -> hanging nodes
  std::vector<int> pptr(n + 1, 0), pdep;
  std::vector<float> pw;
  for (int j = 0; j < (nd ? dptr[nd] : 0); j++) pptr[dconn[j] + 1]++;
  for (int i = 0; i < n; i++) pptr[i + 1] += pptr[i];
  pdep.resize(pptr[n]);
  pw.resize(pptr[n]);
  {
    std::vector<int> fill(pptr.begin(), pptr.end() - 1);
    for (int d = 0; d < nd; d++)
      for (int j = dptr[d]; j < dptr[d + 1]; j++) {
        pdep[fill[dconn[j]]] = d;
        pw[fill[dconn[j]]++] = (float)dw[j];
      }
  }
  // tiles never straddle two independent components
  std::vector<int> tstart;
  L.seg_tile.assign(1, 0);
  for (size_t sg = 0; sg + 1 < L.seg_node0.size(); sg++) {
    for (int a = L.seg_node0[sg]; a < L.seg_node0[sg + 1]; a += L.tile_nodes) tstart.push_back(a);
    L.seg_tile.push_back((int)tstart.size());
  }
  tstart.push_back(n);
  const int ntiles = (int)tstart.size() - 1;
  std::vector<int> node0(ntiles + 1), ext_ptr{0}, ext, nacc(ntiles), el_ptr{0};
  std::vector<TileElem> els, mel;
  std::vector<int> mb_ptr((size_t)ntiles * (MFMA_WAVES + 1), 0);
  std::vector<int> comb_ptr(n + 1, 0), comb_slot;
  std::vector<float> comb_w;
  std::vector<int> stamp_n(n + nd, -1), slot(n + nd, 0), stamp_e(ne, -1);
  std::vector<uint32_t> col;
  int max_slots = 0, max_acc = 0, ncolors = 0, max_mrec = 0, max_rec = 0;
  bool single = true;
  for (int e = 0; e < ne; e++)
    if (L.cls[e] != L.cls[0]) single = false;
  std::vector<int> tel, rdeps, rdep_wave;
  for (int t = 0; t < ntiles; t++) {
    const int n0 = tstart[t], n1 = std::min(tstart[t + 1], n0 + L.tile_nodes), nown = n1 - n0;
    node0[t] = n0;
    rdeps.clear();
    rdep_wave.clear();
    tel.clear();
    const int qw = (nown + MFMA_WAVES - 1) / MFMA_WAVES;  // owned nodes per wave
    for (int p = n0; p < n1; p++) {
      stamp_n[p] = t;
      slot[p] = p - n0;
      for (int j = pptr[p]; j < pptr[p + 1]; j++) {
        const int x = n + pdep[j];
        if (stamp_n[x] != t) {
          stamp_n[x] = t;
          slot[x] = nown + (int)rdeps.size();
          rdeps.push_back(pdep[j]);
          rdep_wave.push_back((p - n0) / qw);  // the wave of its first owned parent
        }
      }
    }
    const int na = nown + (int)rdeps.size();
    auto add_elems = [&](int x) {
      for (int j = iptr[x]; j < iptr[x + 1]; j++)
        if (stamp_e[iel[j]] != t) {
          stamp_e[iel[j]] = t;
          tel.push_back(iel[j]);
        }
    };
    for (int p = n0; p < n1; p++) add_elems(p);
    for (int d : rdeps) add_elems(n + d);
    std::sort(tel.begin(), tel.end());
    for (int d : rdeps) ext.push_back(-(d + 1));
    int ns = na;
    const size_t el0 = els.size();
    for (int e : tel) {
      TileElem te{};
      te.elem = e;
      te.cls = (uint8_t)L.cls[e];
      for (int a = 0; a < 8; a++) {
        const int g = L.conn[8 * e + a];
        const int x = xid(g);
        if (stamp_n[x] != t) {
          stamp_n[x] = t;
          slot[x] = ns++;
          ext.push_back(g);
        }
        if (slot[x] >= 65528) {
          last_error = "problem: tile slot index overflow";
          return -1;
        }
        te.lc[a] = (uint16_t)tile_pos(slot[x]);
        if (slot[x] < na) te.mask |= (uint8_t)(1u << a);
      }
      els.push_back(te);
    }
    // group equal corner masks (fewer divergent row blocks per wave), then
    // colour greedily over the accumulated slots
    std::stable_sort(els.begin() + el0, els.end(), [](const TileElem &a, const TileElem &b) {
      return a.mask > b.mask;
    });
    col.assign(round8(na), 0u);
    for (size_t k = el0; k < els.size(); k++) {
      uint32_t used = 0;
      for (int a = 0; a < 8; a++)
        if (els[k].mask & (1u << a)) used |= col[els[k].lc[a]];
      int c = 0;
      while (used & (1u << c)) c++;
      if (c >= MAX_COLORS) {
        last_error = "problem: tile needs too many colours";
        return -1;
      }
      els[k].color = (uint8_t)c;
      ncolors = std::max(ncolors, c + 1);
      for (int a = 0; a < 8; a++)
        if (els[k].mask & (1u << a)) col[els[k].lc[a]] |= 1u << c;
    }
    // scatter phases follow the colours: sort by colour, then corner mask
    std::stable_sort(els.begin() + el0, els.end(), [](const TileElem &a, const TileElem &b) {
      return a.color != b.color ? a.color < b.color : a.mask > b.mask;
    });
    // matrix-core batches: each of the MFMA_WAVES waves owns a contiguous
    // share of the owned nodes (and the hanging nodes whose first owned
    // parent it owns) and takes every element touching them, with the rows
    // of those corners only; per wave, records are coloured over its
    // corners and (colour, class) groups are padded to 16 records
    {
      auto wave_of = [&](int pos) {
        const int sl = tile_pos(pos);  // tile_pos is an involution
        return sl < nown ? sl / qw : (sl < na ? rdep_wave[sl - nown] : -1);
      };
      int *cp = &mb_ptr[(size_t)t * (MFMA_WAVES + 1)];
      std::vector<TileElem> grp;
      for (int w = 0; w < MFMA_WAVES; w++) {
        cp[w] = (int)(mel.size() / 16);
        grp.clear();
        for (size_t k = el0; k < els.size(); k++) {
          TileElem r = els[k];
          uint8_t m = 0;
          for (int a = 0; a < 8; a++)
            if ((r.mask & (1u << a)) && wave_of(r.lc[a]) == w) m |= (uint8_t)(1u << a);
          if (!m) continue;
          r.mask = m;
          grp.push_back(r);
        }
        col.assign(round8(na), 0u);
        for (auto &r : grp) {
          uint32_t used = 0;
          for (int a = 0; a < 8; a++)
            if (r.mask & (1u << a)) used |= col[r.lc[a]];
          int c = 0;
          while (used & (1u << c)) c++;
          if (c >= MAX_COLORS) {
            last_error = "problem: tile needs too many colours";
            return -1;
          }
          r.color = (uint8_t)c;
          for (int a = 0; a < 8; a++)
            if (r.mask & (1u << a)) col[r.lc[a]] |= 1u << c;
        }
        std::stable_sort(grp.begin(), grp.end(), [](const TileElem &a, const TileElem &b) {
          return a.color != b.color ? a.color < b.color : a.cls < b.cls;
        });
        for (size_t k = 0; k < grp.size();) {
          const uint8_t c = grp[k].color, cl = grp[k].cls;
          size_t k1 = k;
          while (k1 < grp.size() && grp[k1].color == c && grp[k1].cls == cl) mel.push_back(grp[k1++]);
          TileElem pad{};
          pad.elem = grp[k].elem;
          pad.color = c;
          pad.cls = cl;
          while (mel.size() % 16) mel.push_back(pad);
          k = k1;
        }
      }
      cp[MFMA_WAVES] = (int)(mel.size() / 16);
    }
    nacc[t] = (int)rdeps.size();
    ext_ptr.push_back((int)ext.size());
    el_ptr.push_back((int)els.size());
    max_slots = std::max(max_slots, ns);
    max_acc = std::max(max_acc, na);
    max_rec = std::max(max_rec, (int)(els.size() - el0));
    {
      const int *cp = &mb_ptr[(size_t)t * (MFMA_WAVES + 1)];
      max_mrec = std::max(max_mrec, 16 * (cp[MFMA_WAVES] - cp[0]));
    }
    for (int p = n0; p < n1; p++) {
      for (int j = pptr[p]; j < pptr[p + 1]; j++) {
        comb_slot.push_back(tile_pos(slot[n + pdep[j]]));
        comb_w.push_back(pw[j]);
      }
      comb_ptr[p + 1] = (int)comb_slot.size();
    }
  }
  node0[ntiles] = n;
  {
    // the schedule must be race-free before it reaches the device
    std::vector<int> na_t(ntiles);
    for (int t = 0; t < ntiles; t++) na_t[t] = node0[t + 1] - node0[t] + nacc[t];
    std::string err;
    if (check_tile_schedule(ntiles, el_ptr.data(), els.data(), na_t.data(), MFMA_WAVES,
                            mb_ptr.data(), mel.data(), err)) {
      last_error = "problem: tile schedule check failed: " + err;
      return -1;
    }
  }
  if (comb_slot.empty()) {
    comb_slot.push_back(0);
    comb_w.push_back(0.0f);
  }
  if (ext.empty()) ext.push_back(0);
  if (L.t_node0.upload(node0) || L.t_ext_ptr.upload(ext_ptr) || L.t_ext.upload(ext) ||
      L.t_nacc.upload(nacc) || L.t_el_ptr.upload(el_ptr) || L.t_el.upload(els) ||
      L.t_comb_ptr.upload(comb_ptr) || L.t_comb_slot.upload(comb_slot) ||
      L.t_comb_w.upload(comb_w) || L.t_mb_ptr.upload(mb_ptr) || L.t_mel.upload(mel)) {
    last_error = "problem: device allocation failed";
    return -1;
  }
  {
    std::vector<int> me(mel.size()), ee(els.size());
    for (size_t k = 0; k < mel.size(); k++) me[k] = mel[k].elem;
    for (size_t k = 0; k < els.size(); k++) ee[k] = els[k].elem;
    if (L.t_mel_elem.upload(me) || L.t_el32.upload(els) || L.t_el_elem.upload(ee)) {
      last_error = "problem: device allocation failed";
      return -1;
    }
  }
  Tiles &T = L.tiles;
  T.n_tiles = ntiles;
  // large tiles take 512 threads on the vector ALUs, so that the LDS they
  // need (4 tiles per CU) still fills the CU with waves
  T.threads = L.tile_nodes >= 768 ? 512 : TILE_THREADS;
  T.single_class = single;
  T.max_slots = max_slots;
  T.max_mrec = max_mrec;
  T.max_rec = max_rec;
  T.max_acc = max_acc;
  T.ncolors = ncolors;
  T.node0 = L.t_node0.p;
  T.ext_ptr = L.t_ext_ptr.p;
  T.ext = L.t_ext.p;
  T.nacc_dep = L.t_nacc.p;
  T.el_ptr = L.t_el_ptr.p;
  T.el = L.t_el.p;
  T.comb_ptr = L.t_comb_ptr.p;
  T.comb_slot = L.t_comb_slot.p;
  T.comb_w = L.t_comb_w.p;
  T.dep_ptr = L.dep_ptr.p;
  T.dep_conn = L.dep_conn.p;
  T.dep_w = L.dep_w.p;
  T.bcmask = L.d_bcmask.p;
  T.mb_ptr = L.t_mb_ptr.p;
  T.mel = L.t_mel.p;
  T.mel_elem = L.t_mel_elem.p;
  T.n_mrec = (int)mel.size();
  T.el32 = L.t_el32.p;
  T.el_elem = L.t_el_elem.p;
  T.n_rec = (int)els.size();
  if (have_design_) {
    launch_rec_scale(T.n_mrec, T.mel_elem, L.s32.p, reinterpret_cast<uint32_t *>(T.mel), st_);
    launch_rec_scale(T.n_rec, T.el_elem, L.s32.p, reinterpret_cast<uint32_t *>(T.el32), st_);
  }
  T.afrag = reinterpret_cast<const uint4 *>(afrag_.p);
  T.afrag32 = afrag32_.p;
  T.use_mfma = true;
  return 0;
}

int Problem::build_transfer(int l) {
  Level &F = levels_[l];
  Level &Cc = levels_[l + 1];
  InterpCSR P;
  int dev = 0;
  (void)hipGetDevice(&dev);
  F.nested = F.forest->same_blocks(Cc.forest);
  if (F.nested ? F.forest->create_interpolation_device(Cc.forest, P, dev)
               : F.forest->create_interpolation_geometric(Cc.forest, P)) {
    last_error = "problem: interpolation " + std::to_string(l) + ": " + F.forest->last_error;
    return -1;
  }
  if (P.nrows != F.n_nodes || P.ncols != Cc.n_nodes) {
    last_error = "problem: interpolation size mismatch";
    return -1;
  }
  std::vector<float> pw(P.w.begin(), P.w.end());
  // rows of 2^e equal weights 2^-e (every row of a uniformly refined region)
  // carry 1 + e instead of being read weight by weight
  std::vector<uint8_t> wexp(P.nrows, 0);
  for (int r = 0; r < P.nrows; r++) {
    const int len = P.rowp[r + 1] - P.rowp[r];
    int e = 0;
    while ((1 << e) < len) e++;
    if (len < 1 || (1 << e) != len || e > 30) continue;
    bool eq = true;
    for (int j = P.rowp[r]; j < P.rowp[r + 1]; j++)
      if (P.w[j] != std::ldexp(1.0, -e)) eq = false;
    if (eq) wexp[r] = (uint8_t)(1 + e);
  }
  // transpose (restriction), rows in coarse order, entries in fine order
  std::vector<int> rptr(Cc.n_nodes + 1, 0);
  for (int c : P.cols) rptr[c + 1]++;
  for (int i = 0; i < Cc.n_nodes; i++) rptr[i + 1] += rptr[i];
  std::vector<int> rcol(P.cols.size());
  std::vector<float> rw(P.cols.size());
  {
    std::vector<int> fill(rptr.begin(), rptr.end() - 1);
    for (int r = 0; r < P.nrows; r++)
      for (int j = P.rowp[r]; j < P.rowp[r + 1]; j++) {
        const int c = P.cols[j];
        rcol[fill[c]] = r;
        rw[fill[c]++] = (float)P.w[j];
      }
  }
  // element containment: each fine element lies in one coarse element.
  // Without nesting, each fine element belongs to the coarse element holding
  // its centre, and a coarse element's material factor is the volume
  // average over the fine elements it holds.
  const auto &fo = F.forest->octants();
  std::vector<int> owner(F.n_elem);
  std::vector<int> cnt(Cc.n_elem + 1, 0);
  if (F.nested) {
    for (int e = 0; e < F.n_elem; e++) {
      const int E = Cc.forest->locate(fo[e]);
      if (E < 0 || Cc.lev[E] > F.lev[e]) {
        last_error = "problem: coarse forest does not cover the fine forest";
        return -1;
      }
      owner[e] = E;
    }
  } else {
    if (F.forest->locate_centres(Cc.forest, owner)) {
      last_error = "problem: " + F.forest->last_error;
      return -1;
    }
    for (int e = 0; e < F.n_elem; e++)
      if (owner[e] < 0) {
        last_error = "problem: coarse forest does not cover the fine forest";
        return -1;
      }
  }
  for (int e = 0; e < F.n_elem; e++) cnt[owner[e] + 1]++;
  for (int i = 0; i < Cc.n_elem; i++) cnt[i + 1] += cnt[i];
  std::vector<int> child(F.n_elem), place(F.n_elem, 0);
  std::vector<float> vw(F.n_elem);
  if (F.nested) {
    const auto &co = Cc.forest->octants();
    std::vector<int> fill(cnt.begin(), cnt.end() - 1);
    for (int e = 0; e < F.n_elem; e++) {
      const int E = owner[e];
      const int dl = F.lev[e] - Cc.lev[E];
      if (dl > 8) {
        last_error = "problem: a fine element lies more than 8 levels below its coarse element";
        return -1;
      }
      // offset of e inside E in units of e's width
      const int sh = MAX_LEVEL - F.lev[e];
      const int ox = (fo[e].x - co[E].x) >> sh, oy = (fo[e].y - co[E].y) >> sh,
                oz = (fo[e].z - co[E].z) >> sh;
      place[fill[E]] = (dl << 24) | (oz << 16) | (oy << 8) | ox;
      child[fill[E]] = e;
      vw[fill[E]++] = (float)std::ldexp(1.0, -3 * dl);
    }
  } else {
    std::vector<double> vol(F.n_elem), tot(Cc.n_elem, 0.0);
    for (int e = 0; e < F.n_elem; e++) {
      double h[3];
      F.forest->element_extent(fo[e], h);
      vol[e] = h[0] * h[1] * h[2];
      tot[owner[e]] += vol[e];
    }
    for (int E = 0; E < Cc.n_elem; E++)
      if (!(tot[E] > 0.0)) {
        last_error = "problem: a coarse element holds no fine element centre (coarse elements "
                     "smaller than the fine ones)";
        return -1;
      }
    std::vector<int> fill(cnt.begin(), cnt.end() - 1);
    for (int e = 0; e < F.n_elem; e++) {
      const int E = owner[e];
      child[fill[E]] = e;
      vw[fill[E]++] = (float)(vol[e] / tot[E]);
    }
  }
  F.sc_ptr_h = cnt;
  F.sc_child_h = child;
  F.place_h = place;
  if (F.P_ptr.upload(P.rowp) || F.P_col.upload(P.cols) || F.P_w.upload(pw) || F.P_wexp.upload(wexp) ||
      F.R_ptr.upload(rptr) || F.R_col.upload(rcol) || F.R_w.upload(rw) ||
      F.sc_ptr.upload(cnt) || F.sc_child.upload(child) || F.sc_w.upload(vw) ||
      F.g_place.upload(place)) {
    last_error = "problem: device allocation failed";
    return -1;
  }
  F.P = {F.n_nodes, F.P_ptr.p, F.P_col.p, F.P_w.p, nullptr, F.P_wexp.p};
  F.R = {Cc.n_nodes, F.R_ptr.p, F.R_col.p, F.R_w.p, nullptr, F.P_wexp.p};
  return 0;
}

int Problem::ensure_vecs(int l, int precision, int nrhs) {
  Level &L = levels_[l];
  const size_t n = (size_t)L.ndof;
  const size_t ns = 3 * (size_t)L.wg_nodes.n;
  auto grow = [&](auto &v) -> int {
    using T = std::remove_reference_t<decltype(*v.b.p)>;
    if (!v.ready) {
      if (v.dinv.alloc(n) || v.partial.alloc(ns)) return -1;
      v.ready = true;
      v.cap = 0;
    }
    if (v.cap >= nrhs) return 0;
    const size_t m = n * (size_t)nrhs;
    if (v.b.alloc(m) || v.x.alloc(m) || v.x2.alloc(m) || v.r.alloc(m) || v.d.alloc(m)) return -1;
    if (hipMemset(v.d.p, 0, m * sizeof(T)) != hipSuccess ||
        hipMemset(v.x.p, 0, m * sizeof(T)) != hipSuccess ||
        hipMemset(v.x2.p, 0, m * sizeof(T)) != hipSuccess)
      return -1;
    v.cap = nrhs;
    gen_++;
    return 0;
  };
  return precision == 1 ? grow(L.v32) : grow(L.v64);
}

// PCG vectors, scalars and partials for nrhs right-hand sides
int Problem::ensure_batch(int nrhs) {
  if (nrhs <= batch_cap_) return 0;
  const Level &L0 = levels_[0];
  const size_t n = (size_t)L0.ndof, m = n * (size_t)nrhs;
  const size_t nidx = (size_t)nrhs * nseg_;
  const size_t np = std::max({(size_t)2 * 2048, (size_t)assemble_num_blocks(L0.n_nodes),
                              (size_t)L0.tiles.n_tiles * nrhs, (size_t)2 * plan_.nblk * nrhs,
                              (size_t)(L0.n_elem + 255) / 256}) + 16;
  if (f_.alloc(m) || u_.alloc(m) || r_.alloc(m) || p_.alloc(m) || q_.alloc(m) || z_.alloc(m) ||
      parts_.alloc(np) || sc_.alloc(std::max(nidx * SC_N, (size_t)SC_N)) || act_.alloc(nidx) ||
      itconv_.alloc(nidx) || thr_.alloc(nidx) || iter_.alloc(1) || nact_.alloc(1)) {
    last_error = "problem: device allocation failed";
    return -1;
  }
  HIPCHK(hipMemset(u_.p, 0, m * sizeof(double)));
  HIPCHK(hipMemset(f_.p, 0, m * sizeof(double)));
  if (!nact_host_) HIPCHK(hipHostMalloc((void **)&nact_host_, 2 * sizeof(int)));
  have_solution_ = false;
  batch_cap_ = nrhs;
  gen_++;
  return 0;
}

// fixed reduction blocks of the finest-level vector, none straddling a
// component
int Problem::build_plan() {
  const Level &L = levels_[0];
  constexpr int CHUNK = 4096;  // dofs per reduction block
  std::vector<int> lo, bseg, sblk{0};
  for (int sg = 0; sg < nseg_; sg++) {
    const int d0 = 3 * L.seg_node0[sg], d1 = 3 * L.seg_node0[sg + 1];
    for (int a = d0; a < d1; a += CHUNK) {
      lo.push_back(a);
      bseg.push_back(sg);
    }
    sblk.push_back((int)lo.size());
  }
  lo.push_back((int)L.ndof);
  if (bseg.empty()) bseg.push_back(0);
  if (pl_blk_lo_.upload(lo) || pl_blk_seg_.upload(bseg) || pl_seg_blk_.upload(sblk) ||
      pl_seg_tile_.upload(L.seg_tile)) {
    last_error = "problem: device allocation failed";
    return -1;
  }
  plan_.nrhs = 1;
  plan_.nseg = nseg_;
  plan_.nblk = (int)lo.size() - 1;
  plan_.ntiles = L.tiles.n_tiles;
  plan_.n = L.ndof;
  plan_.blk_lo = pl_blk_lo_.p;
  plan_.blk_seg = pl_blk_seg_.p;
  plan_.seg_blk = pl_seg_blk_.p;
  plan_.seg_tile = pl_seg_tile_.p;
  return 0;
}

// ---------------------------------------------------------------------------
int Problem::check_design_values(const double *v, int64_t n, bool on_device, const char *what) {
  // SIMP rho^q is defined for rho >= 0; RAMP rho / (1 + q (1 - rho)) needs
  // rho < 1 + 1/q (TMROctConstitutive, src/topology/TMROctConstitutive.cpp:541-546)
  const double lo = 0.0;
  const double hi = mat_.penalty_type == 1 && mat_.q > 0.0
                        ? (1.0 + 1.0 / mat_.q) * (1.0 - 1e-12)
                        : std::numeric_limits<double>::max();
  bool bad = false;
  if (!on_device) {
    for (int64_t i = 0; i < n && !bad; i++) bad = !(v[i] >= lo && v[i] <= hi);
  } else {
    if (chk_.n == 0 && chk_.alloc(1)) {
      last_error = std::string(what) + ": device allocation failed";
      return -1;
    }
    int h = 0;
    HIPCHK(hipMemsetAsync(chk_.p, 0, sizeof(int), st_));
    launch_check_range(n, v, lo, hi, chk_.p, st_);
    HIPCHK(hipMemcpyAsync(&h, chk_.p, sizeof(int), hipMemcpyDeviceToHost, st_));
    HIPCHK(hipStreamSynchronize(st_));
    bad = h != 0;
  }
  if (bad) {
    last_error = std::string(what) +
                 (mat_.penalty_type == 1 ? ": design values must be finite and in [0, 1 + 1/q)"
                                         : ": design values must be finite and non-negative");
    return -1;
  }
  return 0;
}

int Problem::set_design(const double *rho, int n, bool on_device) {
  if (levels_.empty()) {
    last_error = "problem: not initialised";
    return -1;
  }
  if (n != levels_[0].n_elem) {
    last_error = "set_design: expected one density per finest-level element";
    return -1;
  }
  if (!rho) {
    last_error = "set_design: null densities";
    return -1;
  }
  if (check_design_values(rho, n, on_device, "set_design")) return -1;
  HIPCHK(hipMemcpyAsync(rho_.p, rho, n * sizeof(double),
                        on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, st_));
  design_mode_ = 0;
  launch_simp_scale(n, rho_.p, levels_[0].d_lev.p, mat_.q, mat_.k0, mat_.penalty_type,
                    levels_[0].s64.p, levels_[0].s32.p, st_);
  for (size_t l = 0; l + 1 < levels_.size(); l++) {
    Level &F = levels_[l], &Cc = levels_[l + 1];
    launch_restrict_scale(Cc.n_elem, F.sc_ptr.p, F.sc_child.p, F.sc_w.p, F.s64.p, F.d_lev.p,
                          Cc.d_lev.p, Cc.s64.p, Cc.s32.p, st_);
  }
  for (auto &L : levels_) {
    launch_rec_scale(L.tiles.n_mrec, L.tiles.mel_elem, L.s32.p,
                     reinterpret_cast<uint32_t *>(L.tiles.mel), st_);
    launch_rec_scale(L.tiles.n_rec, L.tiles.el_elem, L.s32.p,
                     reinterpret_cast<uint32_t *>(L.tiles.el32), st_);
  }
  have_design_ = true;
  have_setup_ = false;
  return 0;
}

// Nodal design variables (TMROctConformTACSTopoCreator: the design forest is
// the analysis forest, src/topology/TMR_TACSTopoCreator.cpp:714-777).  Every
// level gets its design by the weighted restriction of the next finer one
// (TMRConformFilter::setDesignVars, src/topology/TMRConformFilter.cpp:154-167)
// and per-Gauss-point stiffness scales (TMROctConstitutive::evalTangentStiffness,
// src/topology/TMROctConstitutive.cpp:498-551).
int Problem::set_design_nodal(const double *x, int n, bool on_device) {
  if (levels_.empty()) {
    last_error = "problem: not initialised";
    return -1;
  }
  if (n != levels_[0].n_nodes) {
    last_error = "set_design_nodal: expected one design value per finest-level node";
    return -1;
  }
  if (!x) {
    last_error = "set_design_nodal: null design values";
    return -1;
  }
  if (check_design_values(x, n, on_device, "set_design_nodal")) return -1;
  for (auto &L : levels_) {
    if (L.xn.n < (size_t)L.n_nodes && L.xn.alloc(L.n_nodes)) goto oom;
    if (L.sg64.n < 8 * (size_t)L.n_elem && (L.sg64.alloc(8 * (size_t)L.n_elem) ||
                                            L.sg32.alloc(8 * (size_t)L.n_elem)))
      goto oom;
  }
  {
    const size_t ng = 8 * (size_t)levels_[0].n_elem;
    if ((rho_gp_.n < ng && rho_gp_.alloc(ng)) || (egp_.n < ng && egp_.alloc(ng)) ||
        (exgp_.n < ng && exgp_.alloc(ng)) ||
        (gtmp_.n < (size_t)levels_[0].ndof && gtmp_.alloc(levels_[0].ndof)))
      goto oom;
    if (nomask_.n < (size_t)levels_[0].n_nodes) {
      std::vector<uint8_t> z(levels_[0].n_nodes, 0);
      if (nomask_.upload(z)) goto oom;
    }
  }
  HIPCHK(hipMemcpyAsync(levels_[0].xn.p, x, n * sizeof(double),
                        on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, st_));
  for (size_t l = 0; l + 1 < levels_.size(); l++)
    launch_weight_restrict(levels_[l].R, levels_[l].xn.p, levels_[l + 1].xn.p, st_);
  for (size_t l = 0; l < levels_.size(); l++) {
    Level &L = levels_[l];
    launch_nodal_scale(L.n_elem, L.d_conn.p, L.dep_ptr.p, L.dep_conn.p, L.dep_w.p, L.xn.p,
                       L.d_lev.p, mat_.q, mat_.k0, mat_.penalty_type,
                       l == 0 ? rho_gp_.p : nullptr, L.sg64.p, L.sg32.p, st_);
  }
  design_mode_ = 1;
  have_design_ = true;
  have_setup_ = false;
  return 0;
oom:
  last_error = "set_design_nodal: device allocation failed";
  return -1;
}

int Problem::design_size() const {
  if (levels_.empty()) return 0;
  return design_mode_ ? levels_[0].n_nodes : levels_[0].n_elem;
}

int Problem::level_design(int l, double *x) const {
  if (design_mode_ != 1 || !have_design_) {
    last_error = "level_design: no nodal design set";
    return -1;
  }
  if (l < 0 || l >= (int)levels_.size()) {
    last_error = "level_design: level out of range";
    return -1;
  }
  const Level &L = levels_[l];
  if (hipStreamSynchronize(st_) != hipSuccess ||
      hipMemcpy(x, L.xn.p, L.n_nodes * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) {
    last_error = "level_design: copy failed";
    return -1;
  }
  return 0;
}

template <class T>
void Problem::apply_level(int l, const T *x, int mode, const AsmArgs &a, int nrhs) {
  Level &L = levels_[l];
  std::pair<hipEvent_t, hipEvent_t> *ev = nullptr;
  if (l == 0 && ((timing_ >> mode) & 1)) {
    if (ev_used_ == ev_pool_.size()) {
      std::pair<hipEvent_t, hipEvent_t> e;
      if (hipEventCreate(&e.first) == hipSuccess && hipEventCreate(&e.second) == hipSuccess)
        ev_pool_.push_back(e);
    }
    if (ev_used_ < ev_pool_.size()) ev = &ev_pool_[ev_used_++];
    ev_prec_mask_ |= (sizeof(T) == 4) ? 1 : 2;
  }
  AsmArgs g = a;
  if (!g.x) g.x = (void *)x;
  if (l > 0 && galerkin_) {
    BSR<T> A = bsr<T>(l);
    A.run = run_;
    launch_bsr_apply<T>(A, x, mode, g, st_, nrhs, L.ndof);
    return;
  }
  Tiles tl = L.tiles;
  tl.run = run_;
  tl.nodal = design_mode_ == 1;
  tl.use_mfma = tl.use_mfma && !fp32_valu_ && fp32_mfma_ && !tl.nodal;
  tl.mfma_f32 = fp32_mfma_in32_;
  if (ev) (void)hipEventRecord(ev->first, st_);
  launch_tile_apply<T>(tl, tl.nodal ? gtab<T>() : rtab<T>(), tl.nodal ? gscale<T>(l) : scale<T>(l),
                       x, mode, g, st_, nrhs, L.ndof);
  if (ev) (void)hipEventRecord(ev->second, st_);
}

int Problem::estimate_eigs(int l, const SolveOpts &o, bool defer) {
  // power iteration on Dinv A (FP64) from a fixed pseudo-random start
  // (drawn on the device),
  // lambda_max ~ |Dinv A x_k| / |x_k| after eig_iters steps (x_{k+1} =
  // Dinv A x_k, no renormalisation: the growth is bounded by lambda^k).
  // With defer, the sums stay on the device (slot l of eig_sc_) until
  // collect_eigs reads every level's at once.
  Level &L = levels_[l];
  auto &v = L.v64;
  const int64_t n = L.ndof;
  const size_t need = SC_N * levels_.size();
  if (eig_sc_.n < need && eig_sc_.alloc(need)) {
    last_error = "setup: device allocation failed";
    return -1;
  }
  launch_start_vector(n, L.d_bcmask.p, 0x2545f4914f6cdd1dULL * (uint64_t)(l + 1), v.x.p, st_);
  double *a = v.x.p, *b = v.r.p;
  const int iters = std::max(1, o.eig_iters);
  for (int it = 0; it < iters; it++) {
    AsmArgs g;
    g.y = b;
    apply_level<double>(l, a, ASM_STORE, g);
    launch_dinv_dots(n, v.dinv.p, b, a, parts_.p, st_);
    std::swap(a, b);
  }
  double *slot = eig_sc_.p + (size_t)SC_N * l;
  launch_finalize(reduce_blocks(n), parts_.p, 5, slot, st_);
  if (defer) return 0;
  double hsc[SC_N];
  HIPCHK(hipMemcpyAsync(hsc, slot, SC_N * sizeof(double), hipMemcpyDeviceToHost, st_));
  HIPCHK(hipStreamSynchronize(st_));
  L.lam_max = std::sqrt(hsc[SC_ZQ] / hsc[SC_RZNEW]);
  return 0;
}

// lambda_max of levels [0, nl) from the sums estimate_eigs left on the device
int Problem::collect_eigs(int nl) {
  if (nl <= 0) return 0;
  std::vector<double> h((size_t)SC_N * nl);
  HIPCHK(hipMemcpyAsync(h.data(), eig_sc_.p, h.size() * sizeof(double), hipMemcpyDeviceToHost, st_));
  HIPCHK(hipStreamSynchronize(st_));
  for (int l = 0; l < nl; l++)
    levels_[l].lam_max = std::sqrt(h[(size_t)SC_N * l + SC_ZQ] / h[(size_t)SC_N * l + SC_RZNEW]);
  return 0;
}

// Blocks that share no vertex, directly or through other blocks, are
// independent problems solved side by side.  Each component must be a
// contiguous range of block indices, so that its nodes (numbered block by
// block) are a contiguous range on every level.  A level with blocks of its
// own (a hierarchy that is not nested) must hold the same components in the
// same order.
static int block_components(const BlockConn &bd, std::vector<int> &seg0, std::string &err) {
  const int nb = bd.num_blocks;
  std::vector<int> parent(nb);
  for (int b = 0; b < nb; b++) parent[b] = b;
  std::function<int(int)> root = [&](int b) {
    while (parent[b] != b) b = parent[b] = parent[parent[b]];
    return b;
  };
  std::vector<int> vfirst(bd.num_nodes, -1);
  for (int b = 0; b < nb; b++)
    for (int k = 0; k < 8; k++) {
      const int v = bd.block_conn[8 * b + k];
      if (vfirst[v] < 0) {
        vfirst[v] = b;
      } else {
        const int ra = root(vfirst[v]), rb = root(b);
        if (ra != rb) parent[std::max(ra, rb)] = std::min(ra, rb);
      }
    }
  seg0.assign(1, 0);
  std::vector<int> comp(nb);
  for (int b = 0; b < nb; b++) comp[b] = root(b);
  for (int b = 1; b < nb; b++) {
    if (comp[b] != comp[b - 1]) {
      for (int c : seg0)
        if (comp[c] == comp[b]) {
          err = "problem: each independent component must be a contiguous block range";
          return -1;
        }
      seg0.push_back(b);
    }
  }
  seg0.push_back(nb);
  return 0;
}

int Problem::find_components() {
  if (block_components(levels_[0].forest->conn_data(), seg_block0_, last_error)) return -1;
  nseg_ = (int)seg_block0_.size() - 1;
  std::vector<int> seg0 = seg_block0_;
  for (size_t l = 0; l < levels_.size(); l++) {
    Level &L = levels_[l];
    Forest *f = L.forest;
    if (l > 0 && !f->same_blocks(levels_[l - 1].forest)) {
      if (block_components(f->conn_data(), seg0, last_error)) return -1;
      if ((int)seg0.size() != nseg_ + 1) {
        last_error = "problem: level " + std::to_string(l) +
                     " holds a different number of independent components";
        return -1;
      }
    }
    std::vector<int> bseg(f->conn_data().num_blocks);
    for (int s = 0; s < nseg_; s++)
      for (int b = seg0[s]; b < seg0[s + 1]; b++) bseg[b] = s;
    const auto &octs = f->octants();
    const auto &conn = f->conn();
    const int n = f->num_owned_nodes();
    std::vector<int> lo(nseg_, n), hi(nseg_, -1);
    for (size_t e = 0; e < octs.size(); e++) {
      const int s = bseg[octs[e].block];
      for (int a = 0; a < 8; a++) {
        const int g = conn[8 * e + a];
        if (g >= 0) {
          lo[s] = std::min(lo[s], g);
          hi[s] = std::max(hi[s], g);
        }
      }
    }
    L.seg_node0.assign(nseg_ + 1, 0);
    for (int s = 0; s < nseg_; s++) {
      if (lo[s] != (s ? hi[s - 1] + 1 : 0)) {
        last_error = "problem: component node ranges are not contiguous";
        return -1;
      }
      L.seg_node0[s] = lo[s];
    }
    L.seg_node0[nseg_] = n;
    if (hi[nseg_ - 1] != n - 1) {
      last_error = "problem: component node ranges do not cover the nodes";
      return -1;
    }
  }
  return 0;
}

int Problem::coarse_factor() {
  // K of the coarsest level (identity rows and columns on clamped dofs) as
  // a BSR matrix assembled on the device from its element matrices (K_E of
  // the Galerkin product, or the scaled element stiffness), then per
  // independent component a dense block, its FP64 Cholesky and the explicit
  // inverse for a one-GEMV coarse solve
  const int last = (int)levels_.size() - 1;
  Level &L = levels_[last];
  const int n = (int)L.ndof;
  const bool nodal = design_mode_ == 1;
  const bool gal = galerkin_ && levels_.size() > 1;
  if (!L.has_bsr && build_bsr(last)) return -1;
  if (!gal) {
    if ((L.KE.n < 576 * (size_t)L.n_elem && L.KE.alloc(576 * (size_t)L.n_elem)) ||
        (nodal && Kgd_.n == 0 && Kgd_.upload(Kgh_))) {
      last_error = "coarse: device allocation failed";
      return -1;
    }
    launch_elem_matrices(L.n_elem, L.d_cls.p, nodal ? L.sg64.p : L.s64.p,
                         nodal ? Kgd_.p : K0d_.p, nodal, L.KE.p, st_);
  }
  launch_bsr_assemble(bsr<double>(last), L.bt_ptr.p, L.bt_code.p, L.bt_w.p, L.bsr_row.p, L.KE.p,
                      L.bsr64.p, st_);
  std::vector<int> col0(n), nb(n);
  std::vector<int64_t> aoff(n);
  // component blocks of the inverse, one after the other
  std::vector<size_t> base(nseg_ + 1, 0);
  size_t wmax = 0, wsmax = 0;
  for (int sg = 0; sg < nseg_; sg++) {
    const size_t m = 3 * (size_t)(L.seg_node0[sg + 1] - L.seg_node0[sg]);
    base[sg + 1] = base[sg] + m * m;
    const size_t mp = (size_t)dense_padded((int)m);
    wmax = std::max(wmax, mp * mp);
    wsmax = std::max(wsmax, dense_workspace((int)m));
  }
  if ((Ainv_.n < base[nseg_] && Ainv_.alloc(base[nseg_])) ||
      (cw_.n < wmax && (cw_.alloc(wmax) || cx_.alloc(wmax))) || (chk_.n == 0 && chk_.alloc(1)) ||
      (cws_.n < wsmax && cws_.alloc(wsmax))) {
    last_error = "coarse: device allocation failed";
    return -1;
  }
  if (wsmax && !blas_ && !(blas_ = dense_blas_create())) {
    last_error = "coarse: rocBLAS handle creation failed";
    return -1;
  }
  HIPCHK(hipMemsetAsync(chk_.p, 0, sizeof(int), st_));
  for (int sg = 0; sg < nseg_; sg++) {
    const int n0 = L.seg_node0[sg], n1 = L.seg_node0[sg + 1];
    const int d0 = 3 * n0, m = 3 * (n1 - n0);
    const int mp = dense_padded(m);
    launch_bsr_to_dense(L.bsr_row.p, L.bsr_col.p, L.bsr64.p, n0, n1, L.bsr_ptr_h[n0],
                        L.bsr_ptr_h[n1], mp, cw_.p, st_);
    if (dense_workspace(m)) {
      if (launch_dense_inverse_big(blas_, cw_.p, cx_.p, cws_.p, m, Ainv_.p + base[sg], chk_.p,
                                   st_)) {
        last_error = "coarse: the super-block factorisation failed";
        return -1;
      }
    } else {
      launch_dense_inverse(cw_.p, cx_.p, m, Ainv_.p + base[sg], chk_.p, st_);
    }
    for (int i = 0; i < m; i++) {
      col0[d0 + i] = d0;
      nb[d0 + i] = m;
      aoff[d0 + i] = (int64_t)(base[sg] + (size_t)i * m);
    }
  }
  int bad = 0;
  HIPCHK(hipMemcpyAsync(&bad, chk_.p, sizeof(int), hipMemcpyDeviceToHost, st_));
  HIPCHK(hipStreamSynchronize(st_));
  if (bad) {
    last_error = "coarse Cholesky: the coarsest-level matrix is not positive definite "
                 "(missing boundary conditions?)";
    return -1;
  }
  if (cb_col0_.upload(col0) || cb_nb_.upload(nb) || cb_aoff_.upload(aoff)) {
    last_error = "coarse: device allocation failed";
    return -1;
  }
  coarse_.n = n;
  coarse_.Ainv = Ainv_.p;
  // FP32 V-cycles read an FP32 copy (half the bytes of every coarse solve)
  coarse_.Ainv32 = nullptr;
  if (setup_prec_pending_ == 1) {
    if (Ainv32_.n < base[nseg_] && Ainv32_.alloc(base[nseg_])) {
      last_error = "coarse: device allocation failed";
      return -1;
    }
    launch_convert<double, float>((int64_t)base[nseg_], Ainv_.p, Ainv32_.p, st_);
    coarse_.Ainv32 = Ainv32_.p;
  }
  coarse_.col0 = cb_col0_.p;
  coarse_.nb = cb_nb_.p;
  coarse_.aoff = cb_aoff_.p;
  return 0;
}

int Problem::setup(const SolveOpts &o) {
  if (!have_design_) {
    last_error = "setup: set the design first";
    return -1;
  }
  {
    // a component with no clamped dof keeps its rigid-body modes: singular
    const Level &F = levels_[0];
    for (int sg = 0; sg < nseg_; sg++) {
      bool any = false;
      for (int i = F.seg_node0[sg]; i < F.seg_node0[sg + 1] && !any; i++) any = F.bcmask[i] != 0;
      if (!any) {
        last_error = "setup: component " + std::to_string(sg) +
                     " has no clamped dof (missing boundary conditions?): the stiffness is singular";
        return -1;
      }
    }
  }
  auto t0 = std::chrono::steady_clock::now();
  const int prec = o.precision ? 1 : 0;
  for (size_t l = 0; l < levels_.size(); l++) {
    if (ensure_vecs((int)l, 0) || (prec == 1 && ensure_vecs((int)l, 1))) {
      last_error = "setup: device allocation failed";
      return -1;
    }
  }
  // a smoothed Galerkin coarsest level applies its assembled P^T A P
  if (galerkin_ && !coarse_direct_ && levels_.size() > 1 && !levels_.back().has_bsr &&
      build_bsr((int)levels_.size() - 1))
    return -1;
  // smoothed levels: all but the coarsest, which is solved directly unless
  // coarse_direct_ is off (then it is smoothed like the others)
  const size_t nsm = coarse_direct_ ? levels_.size() - 1 : levels_.size();
  // SOR smooths the assembled matrices: patterns and colourings first, so
  // that galerkin_setup fills the Galerkin levels' values
  if (sor_)
    for (size_t l = 0; l < nsm; l++)
      if (!levels_[l].has_sor && build_sor((int)l)) return -1;
  if (galerkin_ && galerkin_setup(prec)) return -1;
  for (size_t l = 0; l < nsm; l++) {
    Level &L = levels_[l];
    if (sor_) {
      if (sor_assemble((int)l, prec)) return -1;
      L.lam_max = 0.0;
      continue;
    }
    if (galerkin_ && l > 0) {
      // Jacobi diagonal of the Galerkin operator
      launch_bsr_dinv<double>(L.n_nodes, L.bsr_diag.p, L.bsr64.p, L.v64.dinv.p, st_);
      if (prec == 1) launch_bsr_dinv<float>(L.n_nodes, L.bsr_diag.p, L.bsr64.p, L.v32.dinv.p, st_);
      if (estimate_eigs((int)l, o, true)) return -1;
      continue;
    }
    // Jacobi diagonal in both precisions
    if (design_mode_ == 1)
      launch_elem_diag<double>(L.eb, G64_.p, L.sg64.p, L.v64.partial.p, st_, true);
    else
      launch_elem_diag<double>(L.eb, K0d_.p, L.s64.p, L.v64.partial.p, st_);
    AsmArgs a;
    a.y = L.v64.dinv.p;
    launch_assemble<double>(L.asmc, L.d_bcmask.p, L.v64.partial.p, ASM_DIAG, a, st_);
    launch_diag_cross(L.n_xd, L.xd_ptr.p, L.xd_node.p, L.xd_code.p, L.xd_w.p, L.d_cls.p,
                      design_mode_ == 1 ? G64_.p : K0d_.p,
                      design_mode_ == 1 ? L.sg64.p : L.s64.p, design_mode_ == 1,
                      L.d_bcmask.p, L.v64.dinv.p, st_);
    if (prec == 1) launch_convert<double, float>(L.ndof, L.v64.dinv.p, L.v32.dinv.p, st_);
    if (estimate_eigs((int)l, o, true)) return -1;
  }
  setup_prec_pending_ = prec;
  if (coarse_direct_) {
    levels_.back().lam_max = 0.0;
    if (coarse_factor()) return -1;
  }
  // one read-back for every level's eigenvalue sums
  if (!sor_ && collect_eigs((int)nsm)) return -1;
  HIPCHK(hipStreamSynchronize(st_));
  have_setup_ = true;
  setup_precision_ = prec;
  gen_++;
  last_setup_ms_ = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return 0;
}

// Chebyshev smoother (degree k) on Dinv A with eigenvalue bounds
// [lower, upper] * lambda_max (TMR's TACSChebyshevSmoother interval,
// src/TMR_RefinementTools.cpp:46-47).  Step k reads x_in and writes x_out
// (tiles read halo values of x, so the update cannot be in place).
template <class T>
void Problem::cheb_step(int l, const T *b, const T *x_in, T *x_out, int k, const SolveOpts &o,
                        int nrhs, int cflags) {
  Level &L = levels_[l];
  LevelVecs<T> &v = vecs<T>(l);
  const double lmax = o.cheb_upper * L.lam_max, lmin = o.cheb_lower * L.lam_max;
  const double theta = 0.5 * (lmax + lmin), delta = 0.5 * (lmax - lmin);
  const double sigma = theta / delta;
  double rho_prev = 1.0 / sigma, ca = 0.0, cc = 1.0 / theta;
  for (int j = 1; j <= k; j++) {
    const double rho = 1.0 / (2.0 * sigma - rho_prev);
    ca = rho * rho_prev;
    cc = 2.0 * rho / delta;
    rho_prev = rho;
  }
  AsmArgs a;
  a.b = b;
  a.x = (void *)x_in;
  a.x_out = x_out;
  a.d = v.d.p;
  a.dinv = v.dinv.p;
  a.ca = ca;
  a.cc = cc;
  a.cflags = k == 0 ? (cflags & CHEB_D_STORE) : cflags;  // step 0: no previous direction
  apply_level<T>(l, x_in, ASM_CHEB, a, nrhs);
}

template <class T>
T *Problem::first_target(int l, T *x, const SolveOpts &o) {
  const int deg = std::max(1, o.cheb_degree);
  const int pre = o.cheb_pre_degree > 0 ? o.cheb_pre_degree : deg;
  // a smoothed coarsest level runs one sweep of deg steps, no post-sweep
  const int updates = l == (int)levels_.size() - 1 ? deg - 1 : pre - 1 + deg;
  return (updates & 1) ? vecs<T>(l).x2.p : x;
}

template <class T>
double Problem::cheb_c0(int l, const SolveOpts &o) const {
  const Level &L = levels_[l];
  const double lmax = o.cheb_upper * L.lam_max, lmin = o.cheb_lower * L.lam_max;
  return 2.0 / (lmax + lmin);
}

template <class T>
void Problem::vcycle(int l, const T *b, T *x, const SolveOpts &o, int nrhs, bool first_done) {
  const int last = (int)levels_.size() - 1;
  if (l == last && coarse_direct_) {
    CoarseBlocks cb = coarse_;
    cb.run = run_;
    launch_dense_solve<T>(cb, b, x, st_, nrhs);
    return;
  }
  if (l == last && sor_) {
    sor_smooth<T>(l, b, x, nrhs, true);
    return;
  }
  if (l == last) {
    // no direct coarse solve (TMR_CreateTACSMg with use_coarse_direct_solve
    // = 0, src/TMR_RefinementTools.cpp:87-102): the level's smoother once,
    // one Chebyshev sweep from zero
    Level &L = levels_[l];
    LevelVecs<T> &v = vecs<T>(l);
    const int deg = std::max(1, o.cheb_degree);
    T *cur = first_target<T>(l, x, o);
    T *oth = cur == x ? v.x2.p : x;
    if (!first_done)
      launch_cheb_first<T>((int)L.ndof, b, v.dinv.p, cheb_c0<T>(l, o), nullptr, cur, st_, nrhs,
                           run_);
    for (int k = 1; k < deg; k++) {
      const int fl = (k == 1 ? CHEB_D_IS_X : CHEB_D_READ) | (k + 1 < deg ? CHEB_D_STORE : 0);
      cheb_step<T>(l, b, cur, oth, k, o, nrhs, fl);
      std::swap(cur, oth);
    }
    return;
  }
  Level &L = levels_[l];
  Level &Lc = levels_[l + 1];
  LevelVecs<T> &v = vecs<T>(l);
  LevelVecs<T> &c = vecs<T>(l + 1);
  if (sor_) {
    // TACSMg with TACSGaussSeidel (TMR_CreateTACSMg, src/TMR_RefinementTools.cpp:
    // 45-52, 93-94): one symmetric SOR sweep before and after the
    // coarse-grid correction
    sor_smooth<T>(l, b, x, nrhs, true);
    AsmArgs a;
    a.b = b;
    a.y = v.r.p;
    apply_level<T>(l, x, ASM_RESID, a, nrhs);
    TransferCSR R = L.R, P = L.P;
    R.run = P.run = run_;
    launch_restrict<T>(R, Lc.d_bcmask.p, v.r.p, c.b.p, st_, nrhs, L.ndof, Lc.ndof, nullptr, 0.0,
                       nullptr);
    vcycle<T>(l + 1, c.b.p, c.x.p, o, nrhs, false);
    launch_prolong_add<T>(P, L.d_bcmask.p, c.x.p, x, st_, nrhs, Lc.ndof, L.ndof);
    sor_smooth<T>(l, b, x, nrhs, false);
    return;
  }
  const int deg = std::max(1, o.cheb_degree);
  const int pre = o.cheb_pre_degree > 0 ? o.cheb_pre_degree : deg;
  // pre - 1 + deg ping-pong updates follow the first step: it goes where
  // the last update then lands in x
  T *cur = first_target<T>(l, x, o);
  T *oth = cur == x ? v.x2.p : x;
  // the first step from a zero guess leaves d = x, so the second step reads
  // x for d; the last step of each sweep stores no direction
  if (!first_done)
    launch_cheb_first<T>((int)L.ndof, b, v.dinv.p, cheb_c0<T>(l, o), nullptr, cur, st_, nrhs, run_);
  for (int k = 1; k < pre; k++) {
    const int fl = (k == 1 ? CHEB_D_IS_X : CHEB_D_READ) | (k + 1 < pre ? CHEB_D_STORE : 0);
    cheb_step<T>(l, b, cur, oth, k, o, nrhs, fl);
    std::swap(cur, oth);
  }
  AsmArgs a;
  a.b = b;
  a.y = v.r.p;
  apply_level<T>(l, cur, ASM_RESID, a, nrhs);
  TransferCSR R = L.R, P = L.P;
  R.run = P.run = run_;
  // the coarse level's first smoothing step rides on the restriction (its
  // spare buffer, where that level's vcycle starts), unless the coarsest
  // level is solved directly
  const bool fuse = l + 1 < last || !coarse_direct_;
  launch_restrict<T>(R, Lc.d_bcmask.p, v.r.p, c.b.p, st_, nrhs, L.ndof, Lc.ndof,
                     fuse ? c.dinv.p : nullptr, fuse ? cheb_c0<T>(l + 1, o) : 0.0,
                     fuse ? first_target<T>(l + 1, c.x.p, o) : nullptr);
  vcycle<T>(l + 1, c.b.p, c.x.p, o, nrhs, fuse);
  launch_prolong_add<T>(P, L.d_bcmask.p, c.x.p, cur, st_, nrhs, Lc.ndof, L.ndof);
  for (int k = 0; k < deg; k++) {
    cheb_step<T>(l, b, cur, oth, k, o, nrhs, CHEB_D_READ | (k + 1 < deg ? CHEB_D_STORE : 0));
    std::swap(cur, oth);
  }
}

// z = M r for every right-hand side: one V-cycle, FP32 (mixed) or FP64
void Problem::precond(int nrhs, const SolveOpts &o) {
  Level &L0 = levels_[0];
  const int64_t m = L0.ndof * nrhs;
  if (o.precision != 0) {
    vcycle<float>(0, L0.v32.b.p, L0.v32.x.p, o, nrhs);
    launch_convert<float, double>(m, L0.v32.x.p, z_.p, st_);
  } else {
    vcycle<double>(0, r_.p, z_.p, o, nrhs);
  }
}

int Problem::solve(const double *f, double *u, bool on_device, const SolveOpts &o,
                   SolveInfo *info) {
  return solve_multi(1, f, u, on_device, o, info);
}

int Problem::solve_multi(int nrhs, const double *f, double *u, bool on_device, const SolveOpts &o,
                         SolveInfo *infos) {
  if (levels_.empty() || !have_design_) {
    last_error = "solve: set the design first";
    return -1;
  }
  if (nrhs < 1) {
    last_error = "solve: need at least one right-hand side";
    return -1;
  }
  if (!f && nrhs > 1) {
    last_error = "solve: null load vectors";
    return -1;
  }
  if (ensure_batch(nrhs)) return -1;
  const int64_t m = levels_[0].ndof * nrhs;
  const bool warm = !o.zero_guess;
  if (warm && !u) {
    last_error = "solve: zero_guess = 0 needs the initial guess in u";
    return -1;
  }
  if (f)
    HIPCHK(hipMemcpyAsync(f_.p, f, m * sizeof(double),
                          on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, st_));
  if (warm)
    HIPCHK(hipMemcpyAsync(u_.p, u, m * sizeof(double),
                          on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, st_));
  const int r = solve_batch(nrhs, f_.p, u_.p, o, infos, has_ubc_, warm);
  if (r < 0) return r;
  if (u) {
    HIPCHK(hipMemcpyAsync(u, u_.p, m * sizeof(double),
                          on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, st_));
    HIPCHK(hipStreamSynchronize(st_));
  }
  have_solution_ = true;
  return r;
}

// one flexible-PCG iteration for every running (right-hand side, component):
// q = A p, alpha, r, stop test, z = M r, beta, x and p; then the running
// count into pinned slot `slot` and the iteration counter advanced.  Kernels
// and one copy only, so that it can be captured.
void Problem::enqueue_iteration(const SegPlan &pl, int nrhs, double *xsol, const SolveOpts &o,
                                int slot) {
  const bool mixed = o.precision != 0;
  float *r32 = mixed ? levels_[0].v32.b.p : nullptr;
  const int nidx = nrhs * nseg_;
  const double rtol2 = o.rtol * o.rtol, atol2 = o.atol * o.atol;
  // (not while launches are being timed: every timed launch does its work)
  run_ = timing_ ? nullptr : nact_.p;
  AsmArgs a;
  a.y = q_.p;
  a.partials = parts_.p;
  apply_level<double>(0, p_.p, ASM_DOT, a, nrhs);
  launch_b_finalize(pl, FIN_PQ, parts_.p, sc_.p, act_.p, thr_.p, iter_.p, itconv_.p, rtol2, atol2,
                    st_);
  // mixed precision: the finest level's first smoothing step rides on the
  // residual update (into the spare buffer where the V-cycle starts)
  Level &L0u = levels_[0];
  const bool fuse0 = mixed && !sor_;
  launch_b_update_xr(pl, r_.p, q_.p, r32, sc_.p, act_.p, parts_.p, st_,
                     fuse0 ? L0u.v32.dinv.p : nullptr, fuse0 ? cheb_c0<float>(0, o) : 0.0,
                     fuse0 ? first_target<float>(0, L0u.v32.x.p, o) : nullptr);
  launch_b_finalize(pl, FIN_RR, parts_.p, sc_.p, act_.p, thr_.p, iter_.p, itconv_.p, rtol2, atol2,
                    st_);
  // count the segments still running (and advance the iteration counter)
  // before the preconditioner: an iteration that converged skips it
  launch_b_tick(iter_.p, act_.p, nidx, nact_.p, st_);
  // z = M r; in mixed precision the FP64 updates read the FP32 result
  // directly (no FP64 copy of z)
  if (mixed) {
    Level &L0 = levels_[0];
    vcycle<float>(0, L0.v32.b.p, L0.v32.x.p, o, nrhs, fuse0);
    launch_b_dot2<float>(pl, L0.v32.x.p, r_.p, q_.p, act_.p, parts_.p, st_);
    launch_b_finalize(pl, FIN_BETA, parts_.p, sc_.p, act_.p, thr_.p, iter_.p, itconv_.p, rtol2,
                      atol2, st_);
    launch_b_update_p<float>(pl, p_.p, L0.v32.x.p, sc_.p, act_.p, xsol, iter_.p, st_);
  } else {
    precond(nrhs, o);
    launch_b_dot2<double>(pl, z_.p, r_.p, q_.p, act_.p, parts_.p, st_);
    launch_b_finalize(pl, FIN_BETA, parts_.p, sc_.p, act_.p, thr_.p, iter_.p, itconv_.p, rtol2,
                      atol2, st_);
    launch_b_update_p<double>(pl, p_.p, z_.p, sc_.p, act_.p, xsol, iter_.p, st_);
  }
  run_ = nullptr;
  (void)hipMemcpyAsync(nact_host_ + slot, nact_.p, sizeof(int), hipMemcpyDeviceToHost, st_);
}

int Problem::solve_batch(int nrhs, double *rhs, double *xsol, const SolveOpts &o,
                         SolveInfo *infos, bool lift, bool warm) {
  if (!have_setup_ || setup_precision_ != (o.precision ? 1 : 0)) {
    if (setup(o)) return -1;
  }
  const bool mixed = o.precision != 0;
  for (size_t l = 0; l < levels_.size(); l++)
    if (ensure_vecs((int)l, mixed ? 1 : 0, nrhs)) {
      last_error = "solve: device allocation failed";
      return -1;
    }
  if (ensure_batch(nrhs)) return -1;
  fp32_valu_ = o.fp32_valu != 0;
  auto t0 = std::chrono::steady_clock::now();
  Level &L0 = levels_[0];
  const int64_t n = L0.ndof, m = n * nrhs;
  SegPlan pl = plan_;
  pl.nrhs = nrhs;
  const int nidx = nrhs * nseg_;
  const double rtol2 = o.rtol * o.rtol, atol2 = o.atol * o.atol;
  float *r32 = mixed ? L0.v32.b.p : nullptr;
  if (warm) {
    // initial guess u0 (the entry contents of xsol): the PCG solves for the
    // correction against b - A u0 from zero, u0 is added afterwards
    if ((w0_.n < (size_t)m && w0_.alloc(m)) || (wrhs_.n < (size_t)m && wrhs_.alloc(m))) {
      last_error = "solve: device allocation failed";
      return -1;
    }
    HIPCHK(hipMemcpyAsync(w0_.p, xsol, m * sizeof(double), hipMemcpyDeviceToDevice, st_));
    launch_zero_clamped(L0.n_nodes, L0.d_bcmask.p, w0_.p, st_, nrhs, n);
  }
  // x = 0; clamped rows of the right-hand side are zero (TACSAssembler::applyBCs)
  launch_fill<double>(m, 0.0, xsol, st_);
  launch_zero_clamped(L0.n_nodes, L0.d_bcmask.p, rhs, st_, nrhs, n);
  if (lift) {
    // prescribed values: the PCG solves for the free dofs against
    // f - K u_bc (K without the identity rows), the values are added after
    if (lrhs_.n < (size_t)m && lrhs_.alloc(m)) {
      last_error = "solve: device allocation failed";
      return -1;
    }
    HIPCHK(hipMemcpyAsync(lrhs_.p, rhs, m * sizeof(double), hipMemcpyDeviceToDevice, st_));
    Tiles tl = L0.tiles;
    tl.bcmask = nomask_.p;
    tl.nodal = design_mode_ == 1;
    tl.use_mfma = false;
    AsmArgs a;
    a.y = q_.p;
    a.x = (void *)ubc_.p;
    launch_tile_apply<double>(tl, tl.nodal ? G64_.p : R64_.p, tl.nodal ? L0.sg64.p : L0.s64.p,
                              ubc_.p, ASM_STORE, a, st_, 1, n);
    launch_axpy_bcast(n, nrhs, -1.0, q_.p, lrhs_.p, st_);
    launch_zero_clamped(L0.n_nodes, L0.d_bcmask.p, lrhs_.p, st_, nrhs, n);
    rhs = lrhs_.p;
  }
  if (warm) {
    AsmArgs a;
    a.b = rhs;
    a.y = wrhs_.p;
    apply_level<double>(0, w0_.p, ASM_RESID, a, nrhs);
    launch_zero_clamped(L0.n_nodes, L0.d_bcmask.p, wrhs_.p, st_, nrhs, n);
    rhs = wrhs_.p;
  }
  HIPCHK(hipMemcpyAsync(r_.p, rhs, m * sizeof(double), hipMemcpyDeviceToDevice, st_));
  launch_b_dot1(pl, r_.p, r_.p, nullptr, parts_.p, st_);
  launch_b_finalize(pl, FIN_BB, parts_.p, sc_.p, act_.p, thr_.p, iter_.p, itconv_.p, rtol2, atol2,
                    st_);
  if (mixed) launch_convert<double, float>(m, r_.p, r32, st_);
  precond(nrhs, o);
  launch_b_dot1(pl, r_.p, z_.p, act_.p, parts_.p, st_);
  launch_b_finalize(pl, FIN_RZ, parts_.p, sc_.p, act_.p, thr_.p, iter_.p, itconv_.p, rtol2, atol2,
                    st_);
  launch_b_copy(pl, z_.p, p_.p, nullptr, st_);
  HIPCHK(hipMemsetAsync(iter_.p, 0, sizeof(int), st_));
  // count the running segments; iteration counter to 1
  launch_b_tick(iter_.p, act_.p, nidx, nact_.p, st_);
  HIPCHK(hipMemcpyAsync(nact_host_, nact_.p, sizeof(int), hipMemcpyDeviceToHost, st_));
  HIPCHK(hipStreamSynchronize(st_));
  int running = nact_host_[0];
  // iterations: replayed graphs (or enqueued bodies), the stop count of
  // iteration k read while iteration k + 1 runs
  const bool graphs = !o.no_graph && !o.monitor && timing_ == 0;
  if (graphs && running > 0) {
    const bool same = g_gen_ == gen_ && g_nrhs_ == nrhs && g_xsol_ == xsol &&
                      g_opts_.precision == o.precision &&
                      g_opts_.fp32_valu == o.fp32_valu && g_opts_.cheb_degree == o.cheb_degree &&
                      g_opts_.cheb_pre_degree == o.cheb_pre_degree &&
                      g_opts_.cheb_lower == o.cheb_lower && g_opts_.cheb_upper == o.cheb_upper &&
                      g_opts_.rtol == o.rtol && g_opts_.atol == o.atol;
    if (!same || !gexec_[0] || !gexec_[1]) {
      drop_graphs();
      if (!cap_st_) HIPCHK(hipStreamCreateWithFlags(&cap_st_, hipStreamNonBlocking));
      hipStream_t keep = st_;
      for (int k = 0; k < 2; k++) {
        hipGraph_t g = nullptr;
        st_ = cap_st_;
        hipError_t e = hipStreamBeginCapture(cap_st_, hipStreamCaptureModeRelaxed);
        if (e == hipSuccess) {
          enqueue_iteration(pl, nrhs, xsol, o, k);
          e = hipStreamEndCapture(cap_st_, &g);
        }
        st_ = keep;
        if (e == hipSuccess) e = take_launch_error();
        if (e == hipSuccess) e = hipGraphInstantiate(&gexec_[k], g, nullptr, nullptr, 0);
        if (g) (void)hipGraphDestroy(g);
        if (e != hipSuccess) {
          drop_graphs();
          last_error = std::string("solve: graph capture failed: ") + hipGetErrorString(e);
          return -1;
        }
      }
      g_gen_ = gen_;
      g_nrhs_ = nrhs;
      g_xsol_ = xsol;
      g_opts_ = o;
    }
  }
  for (int k = 0; k < 2; k++)
    if (!gev_[k]) HIPCHK(hipEventCreateWithFlags(&gev_[k], hipEventDisableTiming));
  int launched = 0;
  while (running > 0 && launched < o.max_iters) {
    const int slot = launched & 1;
    if (graphs)
      HIPCHK(hipGraphLaunch(gexec_[slot], st_));
    else
      enqueue_iteration(pl, nrhs, xsol, o, slot);
    HIPCHK(hipEventRecord(gev_[slot], st_));
    launched++;
    if (o.monitor || launched >= 2) {
      const int prev = o.monitor ? slot : (slot ^ 1);
      HIPCHK(hipEventSynchronize(gev_[prev]));
      running = nact_host_[prev];
      if (o.monitor) fprintf(stderr, "[tmrx] pcg %3d  running %d of %d\n", launched, running, nidx);
    }
  }
  HIPCHK(hipStreamSynchronize(st_));
  // true residual with the FP64 operator, per component
  {
    AsmArgs a;
    a.b = rhs;
    a.y = q_.p;
    apply_level<double>(0, xsol, ASM_RESID, a, nrhs);
    launch_b_dot1(pl, q_.p, q_.p, nullptr, parts_.p, st_);
    launch_b_finalize(pl, FIN_TRUE, parts_.p, sc_.p, act_.p, thr_.p, iter_.p, itconv_.p, rtol2,
                      atol2, st_);
  }
  if (warm)
    for (int r = 0; r < nrhs; r++) launch_axpy_bcast(n, 1, 1.0, w0_.p + r * n, xsol + r * n, st_);
  if (lift) launch_axpy_bcast(n, nrhs, 1.0, ubc_.p, xsol, st_);
  seg_sc_.resize((size_t)nidx * SC_N);
  seg_itc_.resize(nidx);
  seg_act_.resize(nidx);
  HIPCHK(hipMemcpyAsync(seg_sc_.data(), sc_.p, seg_sc_.size() * sizeof(double),
                        hipMemcpyDeviceToHost, st_));
  HIPCHK(hipMemcpyAsync(seg_itc_.data(), itconv_.p, nidx * sizeof(int), hipMemcpyDeviceToHost, st_));
  HIPCHK(hipMemcpyAsync(seg_act_.data(), act_.p, nidx * sizeof(int), hipMemcpyDeviceToHost, st_));
  HIPCHK(hipStreamSynchronize(st_));
  last_nrhs_ = nrhs;
  const double ms =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  bool all_conv = true;
  for (int r = 0; r < nrhs; r++) {
    double bb = 0, rr = 0, trr = 0;
    int iters = 0;
    bool conv = true;
    for (int sg = 0; sg < nseg_; sg++) {
      const double *h = &seg_sc_[(size_t)(r * nseg_ + sg) * SC_N];
      const int itc = seg_itc_[r * nseg_ + sg];
      bb += h[SC_BB];
      rr += h[SC_RR];
      trr += h[SC_TRR];
      iters = std::max(iters, std::abs(itc));
      if (seg_act_[r * nseg_ + sg] || itc < 0) conv = false;
    }
    if (r == 0) last_true_residual = bb > 0 ? std::sqrt(trr / bb) : 0.0;
    if (infos) {
      SolveInfo &in = infos[r];
      in.iters = iters;
      in.converged = conv ? 1 : 0;
      in.rnorm0 = std::sqrt(bb);
      in.rnorm = std::sqrt(std::max(rr, 0.0));
      in.setup_ms = last_setup_ms_;
      in.solve_ms = ms;
    }
    all_conv = all_conv && conv;
  }
  if (!all_conv) {
    last_error = "solve: PCG did not converge to the requested tolerance";
    return 1;
  }
  return 0;
}

int Problem::segment_info(int rhs, int seg, int *iters, int *converged, double *rel_res) const {
  if (rhs < 0 || rhs >= last_nrhs_ || seg < 0 || seg >= nseg_) return -1;
  const int k = rhs * nseg_ + seg;
  const double *h = &seg_sc_[(size_t)k * SC_N];
  if (iters) *iters = std::abs(seg_itc_[k]);
  if (converged) *converged = (!seg_act_[k] && seg_itc_[k] >= 0) ? 1 : 0;
  if (rel_res) *rel_res = h[SC_BB] > 0 ? std::sqrt(h[SC_TRR] / h[SC_BB]) : 0.0;
  return 0;
}

// Installs a stored load case (load f, displacement u) as the state that
// compliance, stress and gradient calls read; the clamped rows of f are
// zeroed as solve() does.
int Problem::set_state(const double *f, const double *u, bool on_device) {
  if (!have_design_) {
    last_error = "set_state: set the design first";
    return -1;
  }
  if (!f || !u) {
    last_error = "set_state: null load or displacement";
    return -1;
  }
  Level &L = levels_[0];
  const int64_t n = L.ndof;
  const hipMemcpyKind k = on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
  HIPCHK(hipMemcpyAsync(f_.p, f, n * sizeof(double), k, st_));
  HIPCHK(hipMemcpyAsync(u_.p, u, n * sizeof(double), k, st_));
  launch_zero_clamped(L.n_nodes, L.d_bcmask.p, f_.p, st_);
  HIPCHK(hipStreamSynchronize(st_));
  have_solution_ = true;
  return 0;
}

int Problem::compliance(double *c) {
  if (!have_solution_) {
    last_error = "compliance: solve first";
    return -1;
  }
  const int64_t n = levels_[0].ndof;
  launch_dot1(n, f_.p, u_.p, parts_.p, st_);
  launch_finalize(reduce_blocks(n), parts_.p, 4, sc_.p, st_);
  double h[SC_N];
  HIPCHK(hipMemcpyAsync(h, sc_.p, SC_N * sizeof(double), hipMemcpyDeviceToHost, st_));
  HIPCHK(hipStreamSynchronize(st_));
  *c = h[SC_BB];
  return 0;
}

int Problem::compliance_grad(double *g, bool on_device) {
  if (!have_solution_) {
    last_error = "compliance_grad: solve first";
    return -1;
  }
  Level &L = levels_[0];
  if (design_mode_ == 1) {
    launch_nodal_energy(L.n_elem, L.d_conn.p, L.dep_ptr.p, L.dep_conn.p, L.dep_w.p, state_mask(),
                        L.d_cls.p, G64_.p, u_.p, u_.p, egp_.p, st_);
    return nodal_gather(nullptr, g, on_device);
  }
  launch_elem_energy(L.n_elem, L.d_conn.p, L.dep_ptr.p, L.dep_conn.p, L.dep_w.p, state_mask(),
                     L.d_cls.p, K0d_.p, u_.p, energy_.p, st_);
  double *out = g;
  DevArray<double> tmp;
  if (!on_device) {
    if (tmp.alloc(L.n_elem)) {
      last_error = "compliance_grad: allocation failed";
      return -1;
    }
    out = tmp.p;
  }
  launch_compliance_grad(L.n_elem, rho_.p, L.d_lev.p, mat_.q, mat_.penalty_type, energy_.p, out,
                         st_);
  if (!on_device)
    HIPCHK(hipMemcpyAsync(g, out, L.n_elem * sizeof(double), hipMemcpyDeviceToHost, st_));
  HIPCHK(hipStreamSynchronize(st_));
  return 0;
}

// Nodal design gradient: per element corner sum_g N_a(g) (ex_g - p'(rho_g)
// 2^-level e_g) with e_g = psi_e^T K_g u_e in egp_ (the nodal branch of
// TMROctConstitutive::addStressDVSens, src/topology/TMROctConstitutive.cpp:557-654),
// gathered onto the independent nodes (hanging corners pass their share to
// their parents, the transpose of the design dependent-node interpolation)
int Problem::nodal_gather(const double *ex_gp, double *g, bool on_device) {
  Level &L = levels_[0];
  if (ensure_vecs(0, 0)) {
    last_error = "nodal gradient: allocation failed";
    return -1;
  }
  launch_nodal_grad(L.eb, L.d_lev.p, rho_gp_.p, mat_.q, mat_.penalty_type, egp_.p, ex_gp,
                    L.v64.partial.p, st_);
  AsmArgs a;
  a.y = gtmp_.p;
  launch_assemble<double>(L.asmc, nomask_.p, L.v64.partial.p, ASM_GATHER, a, st_);
  double *out = g;
  DevArray<double> tmp;
  if (!on_device) {
    if (tmp.alloc(L.n_nodes)) {
      last_error = "nodal gradient: allocation failed";
      return -1;
    }
    out = tmp.p;
  }
  launch_take_comp0(L.n_nodes, gtmp_.p, out, st_);
  if (!on_device)
    HIPCHK(hipMemcpyAsync(g, out, L.n_nodes * sizeof(double), hipMemcpyDeviceToHost, st_));
  HIPCHK(hipStreamSynchronize(st_));
  return 0;
}

StressArgs Problem_stress_args(const Material &m, double P, const double *CB, const double *vol) {
  StressArgs sa;
  sa.P = P;
  sa.eps = m.eps;
  sa.ys = m.ys;
  sa.CB = CB;
  sa.vol = vol;
  return sa;
}

int Problem::stress_pnorm(double P, double *f) {
  if (!have_solution_) {
    last_error = "stress_pnorm: solve first";
    return -1;
  }
  if (!(P >= 2.0)) {
    last_error = "stress_pnorm: the exponent must be >= 2";
    return -1;
  }
  Level &L = levels_[0];
  const StressArgs sa = Problem_stress_args(mat_, P, CB_.p, cvol_.p);
  const int nb = (L.n_elem + 255) / 256;
  if ((size_t)nb > parts_.n) {
    last_error = "stress_pnorm: reduction buffer too small";
    return -1;
  }
  launch_stress_sum(L.n_elem, L.d_conn.p, L.dep_ptr.p, L.dep_conn.p, L.dep_w.p, state_mask(),
                    L.d_cls.p, L.d_lev.p, rho_.p, design_mode_ == 1 ? rho_gp_.p : nullptr, u_.p,
                    sa, parts_.p, st_);
  launch_finalize(nb, parts_.p, 4, sc_.p, st_);
  double h[SC_N];
  HIPCHK(hipMemcpyAsync(h, sc_.p, SC_N * sizeof(double), hipMemcpyDeviceToHost, st_));
  HIPCHK(hipStreamSynchronize(st_));
  *f = std::pow(h[SC_BB], 1.0 / P);
  last_stress_sum_ = h[SC_BB];
  return 0;
}

int Problem::stress_pnorm_grad(double P, double *g, bool on_device, const SolveOpts &o,
                               SolveInfo *info) {
  double fval;
  if (stress_pnorm(P, &fval)) return -1;
  Level &L = levels_[0];
  const StressArgs sa = Problem_stress_args(mat_, P, CB_.p, cvol_.p);
  // coef = S^(1/P - 1) into the scalar slot the gradient kernel reads
  const double coef = std::pow(last_stress_sum_, 1.0 / P - 1.0);
  HIPCHK(hipMemcpyAsync(sc_.p + SC_ZQ, &coef, sizeof(double), hipMemcpyHostToDevice, st_));
  const bool nodal = design_mode_ == 1;
  ElemBatches eb = L.eb;
  eb.bcmask = state_mask();
  launch_stress_grad(eb, L.d_lev.p, L.d_cls.p, rho_.p, nodal ? rho_gp_.p : nullptr, u_.p, sa,
                     sc_.p, L.v64.partial.p, nodal ? exgp_.p : dfdrho_.p, st_);
  AsmArgs a;
  a.y = adjf_.p;
  launch_assemble<double>(L.asmc, L.d_bcmask.p, L.v64.partial.p, ASM_GATHER, a, st_);
  HIPCHK(hipStreamSynchronize(st_));
  // adjoint K psi = df/du (K symmetric); u_ and f_ are left untouched
  const double tr = last_true_residual;
  SolveOpts oa = o;
  oa.zero_guess = 1;  // the adjoint starts from zero
  int r = solve_batch(1, adjf_.p, psi_.p, oa, info);
  last_adjoint_residual = last_true_residual;
  last_true_residual = tr;
  if (r != 0) return r < 0 ? -1 : 1;
  if (nodal) {
    launch_nodal_energy(L.n_elem, L.d_conn.p, L.dep_ptr.p, L.dep_conn.p, L.dep_w.p, state_mask(),
                        L.d_cls.p, G64_.p, u_.p, psi_.p, egp_.p, st_);
    return nodal_gather(exgp_.p, g, on_device);
  }
  launch_elem_energy2(L.n_elem, L.d_conn.p, L.dep_ptr.p, L.dep_conn.p, L.dep_w.p, state_mask(),
                      L.d_cls.p, K0d_.p, u_.p, psi_.p, energy_.p, st_);
  double *out = g;
  DevArray<double> tmp;
  if (!on_device) {
    if (tmp.alloc(L.n_elem)) {
      last_error = "stress_pnorm_grad: allocation failed";
      return -1;
    }
    out = tmp.p;
  }
  launch_adjoint_grad(L.n_elem, rho_.p, L.d_lev.p, mat_.q, mat_.penalty_type, energy_.p,
                      dfdrho_.p, out, st_);
  if (!on_device)
    HIPCHK(hipMemcpyAsync(g, out, L.n_elem * sizeof(double), hipMemcpyDeviceToHost, st_));
  HIPCHK(hipStreamSynchronize(st_));
  return 0;
}

// One component of the V-cycle on its own, in the precision of the
// V-cycle (o.precision): the pieces that the solve only shows through its
// iteration count, so that each can be compared with an assembled operator.
// The input is level lin's vector, the output level lout's.
template <class T>
int Problem::mg_op_t(int op, int l, const double *x, double *y, bool on_device,
                     const SolveOpts &o) {
  const int lin = op == MG_PROLONG ? l + 1 : l;
  const int lout = op == MG_RESTRICT ? l + 1 : l;
  Level &Li = levels_[lin], &Lo = levels_[lout];
  LevelVecs<T> &vi = vecs<T>(lin), &vo = vecs<T>(lout);
  const bool f32 = sizeof(T) == 4;
  // the input goes to the b buffer of its level (x for the prolongation,
  // which reads the coarse level's solution); FP32 through the p_ staging
  T *in = op == MG_PROLONG ? vi.x.p : vi.b.p;
  const hipMemcpyKind kin = on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
  if (f32) {
    HIPCHK(hipMemcpyAsync(p_.p, x, Li.ndof * sizeof(double), kin, st_));
    launch_convert<double, float>(Li.ndof, p_.p, reinterpret_cast<float *>(in), st_);
  } else {
    HIPCHK(hipMemcpyAsync(in, x, Li.ndof * sizeof(double), kin, st_));
  }
  T *out = nullptr;
  switch (op) {
    case MG_APPLY: {
      AsmArgs a;
      a.y = vo.r.p;
      apply_level<T>(l, in, ASM_STORE, a, 1);
      out = vo.r.p;
      break;
    }
    case MG_RESTRICT:
      launch_restrict<T>(Li.R, Lo.d_bcmask.p, in, vo.b.p, st_, 1, Li.ndof, Lo.ndof);
      out = vo.b.p;
      break;
    case MG_PROLONG:
      HIPCHK(hipMemsetAsync(vo.x.p, 0, Lo.ndof * sizeof(T), st_));
      launch_prolong_add<T>(Lo.P, Lo.d_bcmask.p, in, vo.x.p, st_, 1, Li.ndof, Lo.ndof);
      out = vo.x.p;
      break;
    case MG_SMOOTH: {
      if (sor_) {  // one symmetric SOR sweep from zero
        sor_smooth<T>(l, in, vo.x.p, 1, true);
        out = vo.x.p;
        break;
      }
      // the V-cycle's pre-smoothing: c0 Dinv b, then pre - 1 ping-pong steps
      const int pre = o.cheb_pre_degree > 0 ? o.cheb_pre_degree : std::max(1, o.cheb_degree);
      T *cur = vo.x.p, *oth = vo.x2.p;
      launch_cheb_first<T>((int)Lo.ndof, in, vo.dinv.p, cheb_c0<T>(l, o), nullptr, cur, st_, 1);
      for (int k = 1; k < pre; k++) {
        const int fl = (k == 1 ? CHEB_D_IS_X : CHEB_D_READ) | (k + 1 < pre ? CHEB_D_STORE : 0);
        cheb_step<T>(l, in, cur, oth, k, o, 1, fl);
        std::swap(cur, oth);
      }
      out = cur;
      break;
    }
    case MG_COARSE:
      // the direct solve, or the coarsest level's smoother once
      if (coarse_direct_)
        launch_dense_solve<T>(coarse_, in, vo.x.p, st_, 1);
      else
        vcycle<T>(l, in, vo.x.p, o, 1);
      out = vo.x.p;
      break;
    default:  // MG_VCYCLE
      vcycle<T>(l, in, vo.x.p, o, 1);
      out = vo.x.p;
      break;
  }
  const hipMemcpyKind kout = on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
  if (f32) {
    launch_convert<float, double>(Lo.ndof, reinterpret_cast<const float *>(out), q_.p, st_);
    HIPCHK(hipMemcpyAsync(y, q_.p, Lo.ndof * sizeof(double), kout, st_));
  } else {
    HIPCHK(hipMemcpyAsync(y, out, Lo.ndof * sizeof(double), kout, st_));
  }
  HIPCHK(hipStreamSynchronize(st_));
  return 0;
}

int Problem::mg_op(int op, int l, const double *x, double *y, bool on_device,
                   const SolveOpts &o, double *lam) {
  const int nl = (int)levels_.size();
  if (!have_design_) {
    last_error = "mg_op: set the design first";
    return -1;
  }
  if (op < MG_APPLY || op > MG_VCYCLE || l < 0 || l >= nl) {
    last_error = "mg_op: invalid operation " + std::to_string(op) + " or level " + std::to_string(l);
    return -1;
  }
  if ((op == MG_RESTRICT || op == MG_PROLONG || op == MG_SMOOTH) && l == nl - 1) {
    last_error = "mg_op: level " + std::to_string(l) + " is the coarsest (no transfer or smoother)";
    return -1;
  }
  if (op == MG_COARSE && l != nl - 1) {
    last_error = "mg_op: the coarse solve is on the coarsest level " + std::to_string(nl - 1);
    return -1;
  }
  if (op == MG_APPLY && galerkin_ && l > 0 && l == nl - 1 && coarse_direct_) {
    last_error = "mg_op: the coarsest Galerkin operator of a direct coarse solve is kept "
                 "factored only";
    return -1;
  }
  if (!x || !y) {
    last_error = "mg_op: null vector";
    return -1;
  }
  const int prec = o.precision ? 1 : 0;
  if (!have_setup_ || setup_precision_ != prec) {
    if (setup(o)) return -1;
  }
  fp32_valu_ = o.fp32_valu != 0;
  if (lam) *lam = levels_[l].lam_max;
  return prec ? mg_op_t<float>(op, l, x, y, on_device, o)
              : mg_op_t<double>(op, l, x, y, on_device, o);
}

int Problem::apply(const double *x, double *y, bool on_device, int precision) {
  if (!have_design_) {
    last_error = "apply: set the design first";
    return -1;
  }
  Level &L = levels_[0];
  const int64_t n = L.ndof;
  if (ensure_vecs(0, 0) || (precision && ensure_vecs(0, 1))) {
    last_error = "apply: allocation failed";
    return -1;
  }
  HIPCHK(hipMemcpyAsync(p_.p, x, n * sizeof(double),
                        on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, st_));
  AsmArgs a;
  if (precision) {
    const bool keep = fp32_valu_;
    fp32_valu_ = precision == 2;
    launch_convert<double, float>(n, p_.p, L.v32.b.p, st_);
    a.y = L.v32.r.p;
    apply_level<float>(0, L.v32.b.p, ASM_STORE, a);
    fp32_valu_ = keep;
    launch_convert<float, double>(n, L.v32.r.p, q_.p, st_);
  } else {
    a.y = q_.p;
    apply_level<double>(0, p_.p, ASM_STORE, a);
  }
  HIPCHK(hipMemcpyAsync(y, q_.p, n * sizeof(double),
                        on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, st_));
  HIPCHK(hipStreamSynchronize(st_));
  return 0;
}

// Consistent load of a constant traction on the faces named `name`
// (TopOptUtils.compute3DTractionLoad, tmr/TopOptUtils.py:241-267, with the
// bilinear face integral of a box face: t * area / 4 per face node; hanging
// nodes pass their share to their parents).
int Problem::traction_load(const char *name, const double tr[3], double *f) {
  Level &L = levels_[0];
  std::vector<Oct> octs;
  L.forest->octs_with_name(name, octs);
  std::memset(f, 0, sizeof(double) * L.ndof);
  const auto &dptr = L.forest->dep_ptr();
  const auto &dconn = L.forest->dep_conn();
  const auto &dw = L.forest->dep_weights();
  for (const Oct &o : octs) {
    const int face = o.info;
    double h[3];
    L.forest->element_extent(o, h);
    const int ax = face / 2;
    const double area = h[(ax + 1) % 3] * h[(ax + 2) % 3];
    const int *c = &L.conn[8 * o.tag];
    for (int k = 0; k < 8; k++) {
      const int bit = (ax == 0) ? (k & 1) : (ax == 1) ? ((k >> 1) & 1) : (k >> 2);
      if (bit != face % 2) continue;
      const int g = c[k];
      if (g >= 0) {
        for (int d = 0; d < 3; d++) f[3 * g + d] += 0.25 * area * tr[d];
      } else {
        for (int j = dptr[-g - 1]; j < dptr[-g]; j++)
          for (int d = 0; d < 3; d++) f[3 * dconn[j] + d] += dw[j] * 0.25 * area * tr[d];
      }
    }
  }
  return 0;
}

// Owner-tile size autotuner: for each candidate size, rebuild the tiles of
// every level and time the finest-level operator (FP32 product on the matrix
// cores 4x, as in a V-cycle's smoothing and residual, plus one FP64 product,
// as in a PCG iteration); keep the fastest.  Needs a design.
int Problem::set_coarse_direct(bool on) {
  if (on != coarse_direct_) {
    coarse_direct_ = on;
    have_setup_ = false;
    drop_graphs();
    gen_++;
  }
  return 0;
}

int Problem::set_tiling(int tile_nodes) {
  if (levels_.empty()) {
    last_error = "set_tiling: not initialised";
    return -1;
  }
  if (tile_nodes < 64 || tile_nodes > 4096) {
    last_error = "set_tiling: tile sizes must lie in [64, 4096]";
    return -1;
  }
  int dev = 0, max_lds = 0;
  HIPCHK(hipGetDevice(&dev));
  HIPCHK(hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
  for (size_t l = 0; l < levels_.size(); l++) {
    Level &L = levels_[l];
    L.tile_nodes = l == 0 ? tile_nodes : coarse_tile_nodes(tile_nodes, L.n_nodes);
    if (build_tiles((int)l)) return -1;
    const Tiles &tl = L.tiles;
    const size_t lds64 = 3 * (size_t)(round8(tl.max_slots) + round8(tl.max_acc)) * sizeof(double);
    const size_t lds32 = 24 * (size_t)round8(tl.max_slots) + 16 * (size_t)round8(tl.max_acc);
    if (std::max(lds64, lds32) + 64 > (size_t)max_lds) {
      last_error = "set_tiling: the tiles of level " + std::to_string(l) + " do not fit in LDS";
      return -1;
    }
  }
  if (build_plan()) return -1;
  const int cap = std::max(1, batch_cap_);
  batch_cap_ = 0;
  if (ensure_batch(cap)) return -1;
  drop_graphs();
  gen_++;
  return 0;
}

int Problem::set_fp32_form(int form) {
  if (form < 0 || form > 2) {
    last_error = "set_fp32_form: form must be 0 (vector ALUs), 1 (matrix cores, bf16 split) or 2 "
                 "(matrix cores, FP32 inputs)";
    return -1;
  }
  fp32_mfma_ = form != 0;
  fp32_mfma_in32_ = form == 2;
  drop_graphs();
  gen_++;
  return 0;
}

int Problem::autotune(const int *cands, int ncand, int reps, int *chosen, double *ms_out) {
  if (!have_design_) {
    last_error = "autotune: set the design first";
    return -1;
  }
  if (ncand < 1 || !cands) {
    last_error = "autotune: no candidate tile sizes";
    return -1;
  }
  for (int c = 0; c < ncand; c++)
    if (cands[c] < 64 || cands[c] > 4096) {
      last_error = "autotune: tile sizes must lie in [64, 4096]";
      return -1;
    }
  reps = std::max(1, reps);
  if (ensure_vecs(0, 0) || ensure_vecs(0, 1)) {
    last_error = "autotune: allocation failed";
    return -1;
  }
  Level &L0 = levels_[0];
  const int64_t n = L0.ndof;
  // inputs of the timed launches in the V-cycle's scratch vectors
  // (Chebyshev steps and residuals need b, x, d and a diagonal, r standing
  // in for it; the values only have to be finite; the set-up Dinv is not
  // touched)
  launch_fill<float>(n, 1.0f, L0.v32.b.p, st_);
  launch_fill<float>(n, 1.0f, L0.v32.x.p, st_);
  launch_fill<float>(n, 0.5f, L0.v32.d.p, st_);
  launch_fill<float>(n, 0.25f, L0.v32.r.p, st_);
  launch_fill<double>(n, 1.0, L0.v64.b.p, st_);
  launch_fill<double>(n, 1.0, L0.v64.x.p, st_);
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  const bool keep_valu = fp32_valu_, keep_mfma = fp32_mfma_, keep_in32 = fp32_mfma_in32_;
  fp32_valu_ = false;
  int dev = 0, max_lds = 0;
  HIPCHK(hipGetDevice(&dev));
  HIPCHK(hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
  int best = -1, best_form = 1;
  double best_ms = 0.0;
  for (int c = 0; c < ncand; c++) {
    bool fits = true;
    for (size_t l = 0; l < levels_.size(); l++) {
      levels_[l].tile_nodes = l == 0 ? cands[c] : coarse_tile_nodes(cands[c], levels_[l].n_nodes);
      if (build_tiles((int)l)) {
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        return -1;
      }
      const Tiles &tl = levels_[l].tiles;
      const size_t lds64 = 3 * (size_t)(round8(tl.max_slots) + round8(tl.max_acc)) * sizeof(double);
      const size_t lds32 = 24 * (size_t)round8(tl.max_slots) + 16 * (size_t)round8(tl.max_acc);
      if (std::max(lds64, lds32) + 64 > (size_t)max_lds) fits = false;
    }
    if (!fits) {  // the tiles would not fit in LDS
      if (ms_out) ms_out[c] = -1.0;
      continue;
    }
    // the finest-level launches of one mixed PCG iteration: three FP32
    // Chebyshev steps and an FP32 residual (V-cycle), an FP64 product (PCG)
    AsmArgs ac, ar, a64;
    ac.b = L0.v32.b.p;
    ac.x_out = L0.v32.x2.p;
    ac.d = L0.v32.d.p;
    ac.dinv = L0.v32.r.p;
    ac.ca = 0.5;
    ac.cc = 0.25;
    ac.cflags = CHEB_D_READ | CHEB_D_STORE;
    ar.b = L0.v32.b.p;
    ar.y = L0.v32.d.p;
    a64.b = L0.v64.b.p;
    a64.y = L0.v64.r.p;
    // ms per such group of launches of one operator form: 0 FP32 on the
    // matrix cores (bf16 three-way split), 1 FP32 on the vector ALUs, 2 FP64,
    // 3 FP32 on the matrix cores with FP32 inputs
    auto time_op = [&](int form) -> double {
      fp32_valu_ = form == 1;
      fp32_mfma_ = form == 0 || form == 3;
      fp32_mfma_in32_ = form == 3;
      auto once = [&]() {
        if (form == 2) {
          apply_level<double>(0, L0.v64.x.p, ASM_RESID, a64);
        } else {
          for (int k = 0; k < 3; k++) apply_level<float>(0, L0.v32.x.p, ASM_CHEB, ac);
          apply_level<float>(0, L0.v32.x.p, ASM_RESID, ar);
        }
      };
      once();  // warm
      (void)hipEventRecord(e0, st_);
      for (int r = 0; r < reps; r++) once();
      (void)hipEventRecord(e1, st_);
      if (hipEventSynchronize(e1) != hipSuccess) return -1.0;
      float ms = 0.0f;
      if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess) return -1.0;
      return ms / reps;
    };
    const double tv = time_op(1), t64 = time_op(2);
    // the matrix-core forms cover element densities only
    const double tm = design_mode_ == 1 ? tv : time_op(0);
    const double tm32 = design_mode_ == 1 ? tv : time_op(3);
    if (tm < 0 || tm32 < 0 || tv < 0 || t64 < 0) {
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      fp32_valu_ = keep_valu;
      fp32_mfma_ = keep_mfma;
      fp32_mfma_in32_ = keep_in32;
      HIPCHK(hipGetLastError());
      last_error = "autotune: timing failed";
      return -1;
    }
    int form = 1;
    double tf = tv;
    if (design_mode_ != 1 && tm < tf) {
      form = 0;
      tf = tm;
    }
    if (design_mode_ != 1 && tm32 < tf) {
      form = 3;
      tf = tm32;
    }
    const double per = tf + t64;
    if (ms_out) ms_out[c] = per;
    if (best < 0 || per < best_ms) {
      best = c;
      best_ms = per;
      best_form = form;
    }
  }
  fp32_valu_ = keep_valu;
  fp32_mfma_ = keep_mfma;
  fp32_mfma_in32_ = keep_in32;
  if (best < 0) {
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    last_error = "autotune: no candidate tile size fits in LDS";
    best = -1;
    for (size_t l = 0; l < levels_.size(); l++) {
      levels_[l].tile_nodes = l == 0 ? TILE_NODES : coarse_tile_nodes(TILE_NODES, levels_[l].n_nodes);
      if (build_tiles((int)l)) return -1;
    }
    if (build_plan()) return -1;
    const int cap0 = std::max(1, batch_cap_);
    batch_cap_ = 0;
    if (ensure_batch(cap0)) return -1;
    return -1;
  }
  for (size_t l = 0; l < levels_.size(); l++) {
    levels_[l].tile_nodes = l == 0 ? cands[best] : coarse_tile_nodes(cands[best], levels_[l].n_nodes);
    if (build_tiles((int)l)) return -1;
  }
  // coarse levels large enough to fill the GPU: their own tile size, by the
  // time of the FP32 product on the vector ALUs at that level
  for (size_t l = 1; l + 1 < levels_.size(); l++) {
    Level &Lc = levels_[l];
    if (Lc.n_nodes < 65536) break;
    if (ensure_vecs((int)l, 1)) return -1;
    launch_fill<float>(Lc.ndof, 1.0f, Lc.v32.b.p, st_);
    const int keep = Lc.tile_nodes;
    int bsz = keep;
    double bms = -1.0;
    AsmArgs ac;
    ac.y = Lc.v32.r.p;
    const bool kv = fp32_valu_;
    fp32_valu_ = true;
    for (int cand : {128, 256, 512, 1024}) {
      if ((Lc.n_nodes + cand - 1) / cand < 256) continue;  // at least one tile per CU
      Lc.tile_nodes = cand;
      if (build_tiles((int)l)) return -1;
      const Tiles &tl = Lc.tiles;
      const size_t lds64 = 3 * (size_t)(round8(tl.max_slots) + round8(tl.max_acc)) * sizeof(double);
      if (lds64 + 64 > (size_t)max_lds) continue;
      apply_level<float>((int)l, Lc.v32.b.p, ASM_STORE, ac);
      (void)hipEventRecord(e0, st_);
      for (int r = 0; r < reps; r++) apply_level<float>((int)l, Lc.v32.b.p, ASM_STORE, ac);
      (void)hipEventRecord(e1, st_);
      float ms = 0.0f;
      if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
        fp32_valu_ = kv;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        last_error = "autotune: timing failed";
        return -1;
      }
      if (bms < 0 || ms < bms) {
        bms = ms;
        bsz = cand;
      }
    }
    fp32_valu_ = kv;
    Lc.tile_nodes = bsz;
    if (build_tiles((int)l)) return -1;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  // tiles moved: reduction plan, partial buffers and captured graphs follow
  if (build_plan()) return -1;
  const int cap = std::max(1, batch_cap_);
  batch_cap_ = 0;
  if (ensure_batch(cap)) return -1;
  gen_++;
  fp32_mfma_ = best_form != 1;
  fp32_mfma_in32_ = best_form == 3;
  if (chosen) *chosen = cands[best];
  return 0;
}

template void Problem::vcycle<float>(int, const float *, float *, const SolveOpts &, int, bool);
template void Problem::vcycle<double>(int, const double *, double *, const SolveOpts &, int, bool);

}  // namespace tmrx
