// Multigrid hierarchy + FP64 flexible PCG on one GPU.
//
// Replaces the TACS objects the reference drives from TMR: TACSAssembler
// (src/TMR_TACSCreator.cpp:332-461), TACSMg with Chebyshev smoothing and a
// coarse direct solve (src/TMR_RefinementTools.cpp:37-107), and the Krylov
// solve, compliance and gradient of TMRTopoProblem
// (src/topology/TMRTopoProblem.cpp:967-1060, 1216-1295).
#pragma once
#include <hip/hip_runtime_api.h>

#include <string>
#include <vector>

#include "forest.h"
#include "kernels.h"

namespace tmrx {

template <class T>
struct DevArray {
  T *p = nullptr;
  size_t n = 0;
  DevArray() = default;
  DevArray(const DevArray &) = delete;
  DevArray &operator=(const DevArray &) = delete;
  DevArray(DevArray &&o) noexcept : p(o.p), n(o.n) {
    o.p = nullptr;
    o.n = 0;
  }
  DevArray &operator=(DevArray &&o) noexcept {
    if (this != &o) {
      release();
      p = o.p;
      n = o.n;
      o.p = nullptr;
      o.n = 0;
    }
    return *this;
  }
  ~DevArray() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  int alloc(size_t count) {
    release();
    if (count == 0) count = 1;
    if (hipMalloc((void **)&p, count * sizeof(T)) != hipSuccess) {
      p = nullptr;
      return -1;
    }
    n = count;
    return 0;
  }
  int upload(const T *h, size_t count) {
    if (alloc(count)) return -1;
    if (count && hipMemcpy(p, h, count * sizeof(T), hipMemcpyHostToDevice) != hipSuccess) return -1;
    return 0;
  }
  int upload(const std::vector<T> &v) { return upload(v.data(), v.size()); }
};

struct Material {
  double E = 1.0, nu = 0.3, ys = 1.0;
  double q = 3.0, k0 = 1e-6, eps = 0.3;
  int penalty_type = 0;  // 0 SIMP, 1 RAMP
};

struct SolveOpts {
  double rtol = 1e-8;
  double atol = 1e-30;
  int max_iters = 200;
  int precision = 1;  // 0: FP64 V-cycle, 1: FP32 V-cycle inside FP64 PCG
  int cheb_degree = 2;
  // Chebyshev interval [lower, upper] * lambda_max(D^-1 A).  TMR's own
  // smoother uses [1/30, 1.1] with degree 3 (src/TMR_RefinementTools.cpp:46-47);
  // for degree 2 inside PCG, [1/6, 1.1] takes a third of the iterations on
  // the 4.2M-hex cantilever (10 against 30-32 at [1/30, 1.1]).
  double cheb_lower = 1.0 / 6.0;
  double cheb_upper = 1.1;
  int eig_iters = 10;
  int monitor = 0;
  int zero_guess = 1;
  int fp32_valu = 0;  // FP32 operator on the vector ALUs (else matrix cores)
  int no_graph = 0;   // enqueue every PCG iteration instead of replaying a HIP graph
  int cheb_pre_degree = 0;  // pre-smoothing degree (0: cheb_degree)
};

// multigrid components that Problem::mg_op runs on their own
enum MgOp {
  MG_APPLY = 0,     // y = A_l x
  MG_RESTRICT = 1,  // y = R_l x (level l to l + 1, clamped rows zero)
  MG_PROLONG = 2,   // y = P_l x (level l + 1 to l, clamped rows zero)
  MG_SMOOTH = 3,    // y = the pre-smoothing sweep of level l from zero, right-hand side x
  MG_COARSE = 4,    // y = the coarsest level's direct solve of x
  MG_VCYCLE = 5,    // y = the V-cycle from level l applied to x
};

struct SolveInfo {
  int iters = 0;
  int converged = 0;
  double rnorm0 = 0.0;
  double rnorm = 0.0;
  double setup_ms = 0.0;
  double solve_ms = 0.0;
};

template <class T>
struct LevelVecs {
  DevArray<T> b, x, x2, r, d, dinv, partial, s;  // b, x, x2, r, d: [cap][ndof]
  bool ready = false;
  int cap = 0;  // right-hand sides the per-rhs vectors hold
};

struct Level {
  Forest *forest = nullptr;
  // the next coarser level is a coarsen() of this one (same blocks); else
  // its transfers come from physical positions (no element containment)
  bool nested = true;
  int n_nodes = 0, n_dep = 0, n_elem = 0;
  int64_t ndof = 0;
  // independent components (separate problems sharing the hierarchy):
  // nodes [seg_node0[s], seg_node0[s+1]) belong to component s, owner
  // tiles [seg_tile[s], seg_tile[s+1]) hold them
  std::vector<int> seg_node0, seg_tile;
  int tile_nodes = TILE_NODES;  // owner-tile size (autotuned)
  // host copies
  std::vector<int16_t> lev;
  std::vector<int> cls;
  std::vector<int> conn;
  std::vector<uint8_t> bcmask;
  // device topology
  DevArray<int> wg_node_ptr, wg_nodes, wg_elem_ptr, wg_class;
  DevArray<uint16_t> lconn;
  DevArray<uint8_t> color, d_bcmask;
  DevArray<int> dep_ptr, dep_conn, asm_ptr, asm_slot, d_conn, d_cls;
  DevArray<float> dep_w, asm_w;
  DevArray<int16_t> d_lev;
  ElemBatches eb;
  AsmCSR asmc;
  // owner tiles (single-pass apply)
  DevArray<int> t_mb_ptr;
  DevArray<TileElem> t_mel;
  DevArray<int> t_mel_elem;
  DevArray<int> t_node0, t_ext_ptr, t_ext, t_nacc, t_el_ptr, t_comb_ptr, t_comb_slot;
  DevArray<TileElem> t_el, t_el32;
  DevArray<int> t_el_elem;
  DevArray<float> t_comb_w;
  Tiles tiles;
  // transfer to the next coarser level (P: rows = this level's nodes)
  DevArray<int> P_ptr, P_col, R_ptr, R_col;
  DevArray<float> P_w, R_w;
  DevArray<uint8_t> P_wexp;
  TransferCSR P, R;
  // element-scale restriction from this level into the next coarser one
  // (CSR over the coarse elements); g_place: where each child sits inside
  // its coarse element (level difference, offsets; see launch_galerkin_elem)
  DevArray<int> sc_ptr, sc_child, g_place;
  DevArray<float> sc_w;
  std::vector<int> sc_ptr_h, sc_child_h, place_h;
  // Galerkin coarse operator of this level (levels >= 1, when enabled): the
  // element matrices K_E, and the BSR matrix with its term lists (levels
  // that are smoothed; the coarsest level is factored from K_E)
  DevArray<double> KE;
  DevArray<int> bsr_ptr, bsr_col, bsr_row, bsr_diag, bt_ptr, bt_code;
  DevArray<double> bt_w, bsr64;
  DevArray<float> bsr32;
  std::vector<int> bsr_ptr_h, bsr_col_h;  // host copies of bsr_ptr, bsr_col
  int nnzb = 0;
  bool has_bsr = false;
  // SOR smoothing: the level's nodes in colour-major order (a colouring of
  // the BSR graph, greedy in node order), colour c in
  // sor_nodes[sor_cptr[c] .. sor_cptr[c + 1])
  DevArray<int> sor_nodes;
  std::vector<int> sor_cptr, sor_nodes_h;
  bool has_sor = false;
  // the BSR matrix with its rows in that order (row k = node sor_nodes[k]),
  // sor_map[k'] = the block of the level's BSR that block k' copies
  DevArray<int> sor_ptr, sor_col, sor_map;
  DevArray<double> sor_val64;
  DevArray<float> sor_val32;
  // diagonal terms of corner pairs that share an independent node
  // (elements with hanging corners): per node i, codes 64 e + 8 a + b and
  // weights 2 Z_ai Z_bi
  int n_xd = 0;
  DevArray<int> xd_ptr, xd_node, xd_code;
  DevArray<double> xd_w;
  // stiffness scales
  DevArray<double> s64;
  DevArray<float> s32;
  // nodal design: design value per independent node, scale per Gauss point
  DevArray<double> xn, sg64;
  DevArray<float> sg32;
  double lam_max = 0.0;
  LevelVecs<float> v32;
  LevelVecs<double> v64;
};

// race check of owner-tile schedules (schedule_check.cpp): 0 if the
// records can accumulate without conflicts, else -1 with the reason in err
int check_tile_schedule(int ntiles, const int *rec_ptr, const TileElem *recs, const int *nacc,
                        int waves, const int *wave_ptr, const TileElem *mrecs, std::string &err);

class Problem {
 public:
  Problem() = default;
  ~Problem();
  int init(int nlevels, Forest **forests, const Material &mat, int nbcs, const char **names,
           const int *dof_masks);
  int set_design(const double *rho, int n, bool on_device);
  // one design value per independent node of the finest level; coarse levels
  // take the weighted restriction (R x) / (R 1)
  int set_design_nodal(const double *x, int n, bool on_device);
  // design mode (0 element densities, 1 nodal) and the length of a gradient
  int design_mode() const { return design_mode_; }
  int design_size() const;
  // nodal design values of level l (nodal mode)
  int level_design(int l, double *x) const;
  int setup(const SolveOpts &o);
  int solve(const double *f, double *u, bool on_device, const SolveOpts &o, SolveInfo *info);
  // nrhs load cases at once: f, u are [nrhs][ndof]; infos[nrhs].  Load case
  // 0 becomes the current state.
  int solve_multi(int nrhs, const double *f, double *u, bool on_device, const SolveOpts &o,
                  SolveInfo *infos);
  // per independent component of the last solve: iterations, converged,
  // true relative residual |f - K u| / |f| over the component's dofs
  int segment_info(int rhs, int seg, int *iters, int *converged, double *rel_res) const;
  int num_segments() const { return nseg_; }
  // owner-tile size autotuner (ms per candidate, chosen size)
  int autotune(const int *cands, int ncand, int reps, int *chosen, double *ms);
  int tile_nodes() const { return levels_.empty() ? 0 : levels_[0].tile_nodes; }
  // owner tiles of level l: count and nominal size
  int tile_info(int l, int *ntiles, int *tile_nodes) const {
    if (l < 0 || l >= (int)levels_.size()) return -1;
    if (ntiles) *ntiles = levels_[l].tiles.n_tiles;
    if (tile_nodes) *tile_nodes = levels_[l].tile_nodes;
    return 0;
  }
  bool fp32_matrix_cores() const { return fp32_mfma_; }
  // FP32 operator form: 0 vector ALUs, 1 matrix cores (bf16 three-way
  // split), 2 matrix cores (FP32 inputs); set_fp32_form overrides the
  // autotuner's choice until the next autotune
  int fp32_form() const { return fp32_mfma_ ? (fp32_mfma_in32_ ? 2 : 1) : 0; }
  int set_fp32_form(int form);
  // owner tiles of tile_nodes at the finest level (coarse levels by the
  // autotuner's rule); overrides the autotuner until the next autotune
  int set_tiling(int tile_nodes);
  // the coarsest level solved directly (default) or smoothed like the others
  int set_coarse_direct(bool on);
  // smoother of every smoothed level: Chebyshev (default) or symmetric
  // point SOR with relaxation omega on the assembled level matrices
  int set_smoother(bool sor, double omega);
  bool sor() const { return sor_; }
  // level l's SOR node order (colour-major) and its colour offsets
  int smoother_order(int l, std::vector<int> &nodes, std::vector<int> &cptr);
  bool coarse_direct() const { return coarse_direct_; }
  int set_state(const double *f, const double *u, bool on_device);
  // multigrid options of TMR_CreateTACSMg (src/TMR_RefinementTools.cpp:37-107):
  // Galerkin coarse operators P^T A P on the coarse levels (else
  // rediscretised coarse elements)
  int set_galerkin(bool on);
  bool galerkin() const { return galerkin_; }
  // the BSR Galerkin operator of level l (1 <= l < nlevels - 1): block row
  // pointers, columns and FP64 values (n + 1, nnzb, 9 nnzb entries)
  int level_matrix(int l, std::vector<int> &ptr, std::vector<int> &col,
                   std::vector<double> &val);
  // prescribed values of the clamped dofs (ubc[ndof], entries of free dofs
  // ignored); solutions take them, the free dofs solve K_ff u_f = f_f - K_fc u_c
  int set_bc_values(const double *ubc, int n, bool on_device);
  int compliance(double *c);
  int compliance_grad(double *g, bool on_device);
  // p-norm (exponent P) of the relaxed von Mises failure of the last solve,
  // and its gradient w.r.t. the element densities through an adjoint solve
  int stress_pnorm(double P, double *f);
  // the same value with the reduction scratch in a fixed file-scope device
  // array (no allocation; calls that overlap on one GPU race)
  int stress_pnorm_noalloc(double P, double *f);
  int stress_pnorm_grad(double P, double *g, bool on_device, const SolveOpts &o,
                        SolveInfo *info);
  int apply(const double *x, double *y, bool on_device, int precision);
  // one multigrid component on its own (MG_* below), level l, from zero
  int mg_op(int op, int l, const double *x, double *y, bool on_device, const SolveOpts &o,
            double *lam);
  int traction_load(const char *name, const double tr[3], double *f);
  int num_dof() const { return levels_.empty() ? 0 : (int)levels_[0].ndof; }
  int num_elements() const { return levels_.empty() ? 0 : levels_[0].n_elem; }
  int num_levels() const { return (int)levels_.size(); }
  const Level &level(int i) const { return levels_[i]; }
  void set_stream(hipStream_t s) { st_ = s; }
  // HIP-event timing of every finest-level element-pass launch
  int kernel_timing(int enable);
  int kernel_timing_result(double *total_ms, int *count, int *precision_mask);
  double last_true_residual = 0.0;
  double last_adjoint_residual = 0.0;
  mutable std::string last_error;

 private:
  int build_level(int l, int nbcs, const char **names, const int *masks);
  int build_transfer(int l);
  int build_tiles(int l);
  int ensure_vecs(int l, int precision, int nrhs = 1);
  int ensure_batch(int nrhs);
  int build_plan();
  // Galerkin: check that the prolongation is element-wise trilinear and
  // build the BSR structures and term lists of the coarse levels
  int build_galerkin();
  int build_bsr(int l);
  int galerkin_setup(int precision);
  // SOR: colouring of level l (BSR pattern built if missing), and the
  // values of the smoothed levels' matrices at setup
  int build_sor(int l);
  int sor_assemble(int l, int precision);
  // symmetric SOR (forward then backward sweep) of level l on A x = b,
  // from x = 0 when from_zero
  template <class T>
  void sor_smooth(int l, const T *b, T *x, int nrhs, bool from_zero);
  template <class T>
  BSR<T> bsr(int l);
  template <class T>
  LevelVecs<T> &vecs(int l);
  template <class T>
  const T *K0();
  template <class T>
  const T *scale(int l);
  template <class T>
  const T *gtab();
  template <class T>
  const T *rtab();
  template <class T>
  const T *gscale(int l);
  int nodal_gather(const double *ex_gp, double *g, bool on_device);
  // design values must be finite, >= 0 and (RAMP) below 1 + 1/q, where the
  // penalty's denominator vanishes; what names the caller in the error
  int check_design_values(const double *v, int64_t n, bool on_device, const char *what);
  template <class T>
  void apply_level(int l, const T *x, int mode, const AsmArgs &a, int nrhs = 1);
  template <class T>
  void cheb_step(int l, const T *b, const T *x_in, T *x_out, int k, const SolveOpts &o,
                 int nrhs = 1, int cflags = CHEB_D_READ | CHEB_D_STORE);
  template <class T>
  int mg_op_t(int op, int l, const double *x, double *y, bool on_device, const SolveOpts &o);
  template <class T>
  // first_done: the first smoothing step from zero (c0 Dinv b) is already
  // in the level's spare buffer (written by the restriction or the PCG update)
  void vcycle(int l, const T *b, T *x, const SolveOpts &o, int nrhs = 1, bool first_done = false);
  template <class T>
  double cheb_c0(int l, const SolveOpts &o) const;
  // buffer of level l that the first smoothing step writes (x or its spare)
  template <class T>
  T *first_target(int l, T *x, const SolveOpts &o);
  void precond(int nrhs, const SolveOpts &o);
  int coarse_factor();
  int find_components();
  // batched FP64 flexible PCG from x = 0 on rhs ([nrhs][ndof], clamped rows
  // zeroed in place); work vectors r_, p_, q_, z_
  // warm: x holds the initial guess on entry (else x = 0)
  int solve_batch(int nrhs, double *rhs, double *x, const SolveOpts &o, SolveInfo *infos,
                  bool lift = false, bool warm = false);
  int estimate_eigs(int l, const SolveOpts &o, bool defer = false);
  int collect_eigs(int nl);

  std::vector<Level> levels_;
  Material mat_;
  int nclasses_ = 0;
  std::vector<double> K0h_;  // per class, full-block element
  DevArray<double> K0d_;
  DevArray<float> K0f_;
  DevArray<uint16_t> afrag_;  // bf16 hi/lo MFMA fragments of K0f_
  DevArray<float> afrag32_;   // FP32 MFMA fragments of K0f_
  // nodal design: per-class gradient tables, per-Gauss-point stiffness
  DevArray<double> R64_;  // reflection-block tables of K0, REFL_TABLE per class
  DevArray<float> R32_;
  std::vector<double> Kgh_;  // [class][8][576]
  DevArray<double> G64_;
  DevArray<float> G32_;
  int design_mode_ = 0;      // 0 element densities, 1 nodal
  bool galerkin_ = false, galerkin_built_ = false;
  bool coarse_direct_ = true;
  bool sor_ = false;
  double omega_ = 1.0;
  DevArray<double> Kgd_;  // Kgh_ on the device (Galerkin with a nodal design)
  DevArray<double> rho_gp_, egp_, exgp_, gtmp_;
  DevArray<uint8_t> nomask_;  // all dofs free (full K, unmasked state loads)
  // non-zero prescribed values on clamped dofs
  DevArray<double> ubc_, lrhs_;
  // warm start: the initial guess and the corrected right-hand side
  DevArray<double> w0_, wrhs_;
  bool has_ubc_ = false;
  const uint8_t *state_mask() const { return has_ubc_ ? nomask_.p : levels_[0].d_bcmask.p; }
  DevArray<double> rho_;
  DevArray<double> Ainv_;
  DevArray<double> cw_, cx_;  // coarse factorisation work (padded matrix, L^-1)
  DevArray<double> cws_;      // workspace of the super-block factorisation
  void *blas_ = nullptr;      // rocBLAS handle of the super-block factorisation
  DevArray<int> cb_col0_, cb_nb_;
  DevArray<int64_t> cb_aoff_;
  CoarseBlocks coarse_;
  DevArray<float> Ainv32_;       // FP32 copy of the coarse inverse (FP32 V-cycles)
  int setup_prec_pending_ = 0;   // precision of the setup in progress
  int nseg_ = 1;                 // independent components
  std::vector<int> seg_block0_;  // [nseg + 1] block ranges of the components
  // PCG state: f_, u_ hold the current load case (first of a batch)
  DevArray<double> f_, u_, r_, p_, q_, z_, parts_, sc_;
  DevArray<double> eig_sc_;  // power-iteration sums, SC_N per level
  int batch_cap_ = 0;  // right-hand sides the PCG vectors hold
  SegPlan plan_;
  DevArray<int> pl_blk_lo_, pl_blk_seg_, pl_seg_blk_, pl_seg_tile_;
  DevArray<int> act_, itconv_, iter_, nact_;
  DevArray<int> chk_;  // flag of the design-value check
  DevArray<double> thr_;
  int *nact_host_ = nullptr;  // pinned, two slots (iterations alternate)
  // PCG iteration bodies captured as HIP graphs (two: they write alternate
  // stop-count slots), valid for one (nrhs, options, buffers, setup)
  hipGraphExec_t gexec_[2] = {nullptr, nullptr};
  hipEvent_t gev_[2] = {nullptr, nullptr};
  hipStream_t cap_st_ = nullptr;
  int g_nrhs_ = 0, g_gen_ = -1;
  double *g_xsol_ = nullptr;
  SolveOpts g_opts_;
  int gen_ = 0;  // bumped whenever buffers move or setup reruns
  // while a PCG iteration is enqueued: the device count of running segments;
  // the operator, transfer and smoother launches return at once when it is 0
  // (the iteration launched after the one that converged)
  const int *run_ = nullptr;
  void drop_graphs();
  void enqueue_iteration(const SegPlan &pl, int nrhs, double *xsol, const SolveOpts &o, int slot);
  int last_nrhs_ = 0;
  std::vector<double> seg_sc_;
  std::vector<int> seg_itc_, seg_act_;
  DevArray<double> energy_, psi_, adjf_, dfdrho_, CB_, cvol_;
  bool have_design_ = false, have_setup_ = false, have_solution_ = false;
  int setup_precision_ = -1;
  double last_setup_ms_ = 0.0;
  double last_stress_sum_ = 0.0;
  hipStream_t st_ = nullptr;
  bool fp32_valu_ = false;  // FP32 operator on the vector ALUs
  bool fp32_mfma_ = true;   // autotuned: FP32 operator on the matrix cores
  bool fp32_mfma_in32_ = false;  // ... with FP32 inputs (else bf16 three-way split)
  int timing_ = 0;  // mask of the timed launch modes (bit m = AsmMode m)
  std::vector<std::pair<hipEvent_t, hipEvent_t>> ev_pool_;
  size_t ev_used_ = 0;
  int ev_prec_mask_ = 0;
};

template <>
BSR<double> Problem::bsr<double>(int l);
template <>
BSR<float> Problem::bsr<float>(int l);

}  // namespace tmrx
