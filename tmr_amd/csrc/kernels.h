// HIP kernels of the hex8 solve path (gfx950) and their host launchers.
//
// Data layout in HBM (one multigrid level):
//   * vectors: node-major, 3 dofs interleaved (u_x, u_y, u_z) per independent
//     node, nodes in forest order (Morton order of the node positions);
//   * elements are processed in batches of up to 256 consecutive (Morton
//     order) elements of one stiffness class; a batch stages the distinct
//     nodes it touches ("slots") in LDS.  Slot ids are global:
//     wg_node_ptr[b] .. wg_node_ptr[b+1] belong to batch b;
//   * element -> slot connectivity is 8 x uint16 local indices per element;
//   * the element pass writes one partial sum per slot; the assembly pass
//     gathers, for every independent node, its slots (weight 1) and the slots
//     of the hanging nodes that depend on it (hanging-node weight), so there
//     are no atomics and the summation order is fixed.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace tmrx {

constexpr int WG_ELEMS = 256;
constexpr int MAX_COLORS = 32;

// launch errors: noted after every launch, reported by the next
// synchronising call of the problem
void note_launch(hipError_t e);
hipError_t take_launch_error();
#define TMRX_LAUNCH_CHECK() ::tmrx::note_launch(hipGetLastError())

struct ElemBatches {
  int n_wg = 0;
  int n_elem = 0;
  int ncolors = 0;
  int max_wg_nodes = 0;
  const int *wg_node_ptr = nullptr;  // [n_wg + 1]
  const int *wg_nodes = nullptr;     // [nslots]: >= 0 node, < 0 -(hanging + 1)
  const int *wg_elem_ptr = nullptr;  // [n_wg + 1]
  const int *wg_class = nullptr;     // [n_wg]
  const uint16_t *lconn = nullptr;   // [n_elem * 8]
  const uint8_t *color = nullptr;    // [n_elem]
  const int *dep_ptr = nullptr;      // hanging-node constraints
  const int *dep_conn = nullptr;
  const float *dep_w = nullptr;
  const uint8_t *bcmask = nullptr;   // [n_nodes], bit d set = dof d clamped
};

// Owner tiles: tile t owns the contiguous independent nodes
// [node0[t], node0[t+1]) and holds every element touching them (or touching a
// hanging node one of them supports).  LDS slots: owned nodes, then the
// hanging nodes whose sums the tile needs ("accumulated"), then halo nodes.
constexpr int TILE_NODES = 512;
constexpr int MFMA_WAVES = 4;  // waves of a matrix-core tile, each owning a share of its nodes
constexpr int TILE_THREADS = 256;

// LDS position of tile slot s: an XOR of higher bits into the low three bits
// (a permutation inside each aligned group of 8), so that the corner slots of
// same-colour elements -- Morton strides of 8, 64, 512 -- spread over banks.
__host__ __device__ inline int tile_pos(int s) {
  return s ^ (((s >> 3) ^ (s >> 6) ^ (s >> 9)) & 7);
}
inline int round8(int n) { return (n + 7) & ~7; }

struct TileElem {       // 24 bytes
  uint16_t lc[8];       // LDS slot of each corner
  int32_t elem;         // global element index (stiffness scale)
  uint8_t mask;         // corners whose slot is accumulated
  uint8_t color;        // conflict-free accumulation phase
  uint8_t cls;          // stiffness class
  uint8_t pad;
};
static_assert(sizeof(TileElem) == 24, "tile element record is 24 bytes");

struct Tiles {
  int n_tiles = 0;
  bool single_class = true;
  int max_slots = 0, max_acc = 0;
  int max_rec = 0, max_mrec = 0;  // records per tile (plain, matrix-core padded)
  int ncolors = 0;
  const int *node0 = nullptr;     // [n_tiles + 1]
  const int *ext_ptr = nullptr;   // [n_tiles + 1] non-owned slots
  const int *ext = nullptr;       // global node (>= 0) or -(hanging + 1)
  const int *nacc_dep = nullptr;  // [n_tiles] accumulated hanging slots
  const int *el_ptr = nullptr;    // [n_tiles + 1]
  const TileElem *el = nullptr;
  // the same records for the FP32 element-density apply, with the element
  // word holding the element's FP32 scale (refreshed with the design)
  TileElem *el32 = nullptr;
  const int *el_elem = nullptr;  // element of each record
  int n_rec = 0;
  const int *comb_ptr = nullptr;  // [n_nodes + 1] owned node <- hanging slots
  const int *comb_slot = nullptr; // local slot in the owning tile
  const float *comb_w = nullptr;
  const int *dep_ptr = nullptr;
  const int *dep_conn = nullptr;
  const float *dep_w = nullptr;
  const uint8_t *bcmask = nullptr;
  // matrix-core (FP32) form: per tile and wave, records of the elements
  // touching the wave's share of the nodes, in batches of 16 of one colour
  // and class (padded with mask-0 records); K0 of every class split into
  // bf16 hi/mid/lo MFMA A fragments
  const int *mb_ptr = nullptr;    // [n_tiles * (MFMA_WAVES + 1)] batch offsets
  TileElem *mel = nullptr;        // [16 * batches]; elem word holds the FP32 scale
  const int *mel_elem = nullptr;  // [16 * batches] element of each record
  int n_mrec = 0;
  const uint4 *afrag = nullptr;   // [n_classes][hi, lo][row block][64 lanes]
  // FP32-input matrix-core form: K0 fragments [n_classes][row block][step][64]
  const float *afrag32 = nullptr;
  bool use_mfma = false;
  bool mfma_f32 = false;  // matrix cores with FP32 inputs (else bf16 three-way split)
  // nodal design: the element stiffness is sum_g s_eg K_g over the 8 Gauss
  // points (K0 holds the gradient tables, s_elem 8 scales per element)
  bool nodal = false;
  // threads per tile of the vector-ALU form (256 or 512)
  int threads = TILE_THREADS;
  // when set, launches return at once while *run == 0 (no segment of a
  // batched solve is still running)
  const int *run = nullptr;
};

// Per stiffness class, the nodal-design element tables: dN_a/dx_d at the 8
// Gauss points of the full-block element (G[24 g + 3 a + d], 192 entries),
// then lambda and mu times the Gauss weight |J| (= volume / 8), then 2 pads.
constexpr int NODAL_TABLE = 196;
// Per stiffness class, the element stiffness in the reflection-symmetric
// basis: 8 blocks of 3 x 3 (row-major), block s coupling the Walsh-Hadamard
// coefficients X_0[s^1], X_1[s^2], X_2[s^4], scaled by 1/64.  The
// element-density tile apply takes this table in place of K0.
constexpr int REFL_TABLE = 72;
// Structural non-zeros of those blocks (bit 3 r + c for row r, column c of
// block s) for a box element of an isotropic material, any aspect ratio:
// the rigid translations X_0[0], X_1[0], X_2[0] span the null rows and
// columns of blocks 1, 2 and 4, and blocks 3, 5 and 6 decouple one of their
// three coefficients.  45 of the 72 entries remain.
constexpr unsigned short REFL_NZ[8] = {0x1FF, 0x1B0, 0x145, 0x11B, 0x01B, 0x155, 0x1B1, 0x1FF};
__host__ __device__ constexpr bool refl_nz(int s, int k) { return (REFL_NZ[s] >> k) & 1; }


struct AsmCSR {
  int n_nodes = 0;
  const int *ptr = nullptr;   // [n_nodes + 1]
  const int *slot = nullptr;  // [nnz]
  const float *w = nullptr;   // [nnz]
};

struct TransferCSR {  // rows = output nodes
  int n_rows = 0;
  const int *ptr = nullptr;
  const int *col = nullptr;
  const float *w = nullptr;
  const int *run = nullptr;  // as Tiles::run
  // per fine node: 1 + e when its prolongation row has 2^e equal weights
  // 2^-e (the weights are then not read), 0 otherwise; indexed by row for
  // P and by column for R = P^T
  const uint8_t *wexp = nullptr;
};

// y-side operation performed by the assembly pass
enum AsmMode {
  ASM_STORE = 0,   // y = A x
  ASM_RESID = 1,   // y = b - A x
  ASM_CHEB = 2,    // r = b - A x; d = a d + c Dinv r; x += d
  ASM_DIAG = 3,    // Dinv = 1 / diag(A)   (partials hold element diagonals)
  ASM_DOT = 4,     // y = A x and block partials of x . y
  ASM_GATHER = 5,  // y = sum of the slot partials (clamped rows zero): Z^T g
};

// Chebyshev-step flags: read the previous direction d, take d = x (the step
// after a first step from a zero guess, where x = d), store the new d
enum { CHEB_D_READ = 1, CHEB_D_IS_X = 2, CHEB_D_STORE = 4 };

struct AsmArgs {
  const void *b = nullptr;
  void *y = nullptr;
  void *x = nullptr;       // input of the product (read) or iterate (CHEB)
  void *x_out = nullptr;   // CHEB iterate output of the tile kernel
  void *d = nullptr;       // Chebyshev direction
  const void *dinv = nullptr;
  double ca = 0.0, cc = 0.0;
  int cflags = CHEB_D_READ | CHEB_D_STORE;  // tile apply, ASM_CHEB
  double *partials = nullptr;
};

// single-pass apply with fused epilogue (modes as AsmMode, except ASM_DIAG);
// K0 is the REFL_TABLE table (element densities) or the NODAL_TABLE table
// (tl.nodal) on the vector ALUs, unused by the matrix-core form
template <class T>
void launch_tile_apply(const Tiles &tl, const T *K0, const T *s_elem, const T *x, int mode,
                       const AsmArgs &args, hipStream_t st, int nrhs = 1, int64_t vs = 0);
template <class T>
void launch_elem_apply(const ElemBatches &eb, const T *K0, const T *s_elem, const T *x,
                       T *partial, hipStream_t st);
template <class T>
void launch_elem_diag(const ElemBatches &eb, const T *K0, const T *s_elem, T *partial,
                      hipStream_t st, bool nodal = false);
// Dinv of the nodes with corner-pair terms: dinv = 1 / (1 / dinv + sum_k
// w_k K_e[a][b]) per direction, K_e = s_e K0[cls] (element design) or
// sum_g s_g K_g (nodal design, K0 the NODAL_TABLE tables); clamped dofs kept
void launch_diag_cross(int n, const int *ptr, const int *node, const int *code, const double *w,
                       const int *cls, const double *K0, const double *s, bool nodal,
                       const uint8_t *bcmask, double *dinv, hipStream_t st);
template <class T>
void launch_assemble(const AsmCSR &a, const uint8_t *bcmask, const T *partial, int mode,
                     const AsmArgs &args, hipStream_t st);
int assemble_num_blocks(int n_nodes);

template <class T>
void launch_prolong_add(const TransferCSR &P, const uint8_t *bcmask_f, const T *xc, T *xf,
                        hipStream_t st, int nrhs = 1, int64_t vsc = 0, int64_t vsf = 0);
// rc = R rf (clamped rows zero); with x1, also the coarse level's first
// Chebyshev step from zero, x1 = c0 Dinv_c rc
template <class T>
void launch_restrict(const TransferCSR &R, const uint8_t *bcmask_c, const T *rf, T *rc,
                     hipStream_t st, int nrhs = 1, int64_t vsf = 0, int64_t vsc = 0,
                     const T *dinv_c = nullptr, double c0 = 0.0, T *x1 = nullptr);
// x = c Dinv b (and d = x when d is not null)
template <class T>
void launch_cheb_first(int ndof, const T *b, const T *dinv, double c, T *d, T *x,
                       hipStream_t st, int nrhs = 1, const int *run = nullptr);
// coarsest-level solve: block-diagonal explicit inverse (one block per
// independent component), per row: first column, block size, row offset
struct CoarseBlocks {
  int n = 0;
  const double *Ainv = nullptr;
  const float *Ainv32 = nullptr;  // the same in FP32 (FP32 V-cycles read it when set)
  const int *col0 = nullptr;
  const int *nb = nullptr;
  const int64_t *aoff = nullptr;
  const int *run = nullptr;  // as Tiles::run
};
template <class T>
void launch_dense_solve(const CoarseBlocks &cb, const T *b, T *x, hipStream_t st, int nrhs = 1);
// rows and columns of the padded dense matrices of launch_dense_inverse
int dense_padded(int m);
// explicit inverse of an SPD matrix by blocked Cholesky on the device: W
// holds the matrix row-major with dense_padded(m) rows and columns (identity
// on the padding) and is overwritten by the factor; X (same size) receives
// L^-1; out (m x m, row-major) receives A^-1; *bad |= 1 on a non-positive
// pivot
void launch_dense_inverse(double *W, double *X, int m, double *out, int *bad, hipStream_t st);
// the same for m of at least 2048 over 256-wide super-blocks with the bulk in
// FP64 GEMMs (rocBLAS handle from dense_blas_create, workspace of
// dense_workspace(m) doubles); returns -1 if m is below that or a call fails
int launch_dense_inverse_big(void *blas, double *W, double *X, double *ws, int m, double *out,
                             int *bad, hipStream_t st);
size_t dense_workspace(int m);
void *dense_blas_create();
void dense_blas_destroy(void *h);
// W (mp x mp, row-major) = the BSR block rows [n0, n1) of a component
// (blocks [k0, k1), block rows brow) as a dense matrix of its 3 (n1 - n0)
// dofs, identity on the padding rows
void launch_bsr_to_dense(const int *brow, const int *col, const double *val, int n0, int n1,
                         int k0, int k1, int mp, double *W, hipStream_t st);

// Galerkin coarse operator of one level: 3 x 3 blocks over the independent
// nodes (block row i = node i), row-major blocks, clamped rows and columns
// identity
template <class T>
struct BSR {
  int n = 0;     // block rows
  int nnzb = 0;  // blocks
  const int *ptr = nullptr;  // [n + 1]
  const int *col = nullptr;  // [nnzb] column node, ascending within a row
  const T *val = nullptr;    // [nnzb][9]
  const uint8_t *bcmask = nullptr;
  const int *run = nullptr;  // as Tiles::run
};
using BSR64 = BSR<double>;
// coarse element matrices K_E = sum_e (Q_e x I3)^T K_e (Q_e x I3) over the
// children e of every coarse element (CSR ptr/child over coarse elements;
// place = level difference << 24 | z offset << 16 | y offset << 8 | x
// offset of the child inside E, in child widths).  mode 0: K_e = s_e
// Ktab[cls_e] (finest level, element densities); 1: K_e = Kin[e]; 2:
// K_e = sum_g s_eg Ktab[cls_e][g] (finest level, nodal design)
void launch_galerkin_elem(int mode, int nE, const int *ptr, const int *child, const int *place,
                          const int *cls, const double *s, const double *Ktab, const double *Kin,
                          double *Kout, hipStream_t st);
// BSR values from the element matrices: block k collects the terms
// [tptr[k], tptr[k+1]) (code 64 E + 8 a + b, weight w: w K_E[a][b]); brow[k]
// is its row
// element matrices of one level from its scales: KE[e] = s_e K0[cls_e]
// (element design, Ktab 576 per class) or sum_g s_eg K_g[cls_e] (nodal
// design, Ktab 8 x 576 per class)
void launch_elem_matrices(int ne, const int *cls, const double *s, const double *Ktab, bool nodal,
                          double *KE, hipStream_t st);
void launch_bsr_assemble(const BSR64 &A, const int *tptr, const int *tcode, const double *tw,
                         const int *brow, const double *KE, double *val, hipStream_t st);
// dinv = 1 / diagonal of the blocks at diag[i]
template <class T>
void launch_bsr_dinv(int n, const int *diag, const double *val, T *dinv, hipStream_t st);
// y = A x with the tile apply's epilogues (ASM_STORE, ASM_RESID, ASM_CHEB)
template <class T>
void launch_bsr_apply(const BSR<T> &A, const T *x, int mode, const AsmArgs &g, hipStream_t st,
                      int nrhs = 1, int64_t vs = 0);

// One point SOR sweep (forward, or BACK in reverse) over the colours of a
// node colouring of A's graph: nodes[cptr[c] .. cptr[c + 1]) hold colour c
// (cptr on the host, ncolors + 1 entries) and A's block row k is node
// nodes[k] (colour-major rows, node-index columns); x updated in place
// out[k] = val[map[k]] per 3 x 3 block
void launch_bsr_gather(int nnzb, const int *map, const double *val, double *out, hipStream_t st);
template <class T>
void launch_sor_sweep(const BSR<T> &A, const int *nodes, const int *cptr, int ncolors, bool back,
                      const T *b, T *x, double omega, hipStream_t st, int nrhs = 1,
                      int64_t vs = 0);

template <class TI, class TO>
void launch_convert(int64_t n, const TI *a, TO *b, hipStream_t st);
template <class T>
void launch_fill(int64_t n, T v, T *a, hipStream_t st);
// v[r][i] += alpha y[i] for every right-hand side r (vectors of n values)
void launch_axpy_bcast(int64_t n, int nrhs, double alpha, const double *y, double *v,
                       hipStream_t st);
// y[r][v] = sum_j w_j x[col_j][v] over CSR rows (nv values per node)
void launch_interp_nodal(int nrows, const int *ptr, const int *col, const double *w, int nv,
                         const double *x, double *y, hipStream_t st);
// power-iteration start vector: 0.5 + a hash-drawn U[0, 1) per free dof,
// 0 on clamped dofs
void launch_start_vector(int64_t n, const uint8_t *bcmask, uint64_t seed, double *x,
                         hipStream_t st);
// *bad |= 1 if any of v[0..n) is not finite or outside [lo, hi]
void launch_check_range(int64_t n, const double *v, double lo, double hi, int *bad,
                        hipStream_t st);
// zero the clamped components of a node-major 3-dof vector
void launch_zero_clamped(int n_nodes, const uint8_t *bcmask, double *v, hipStream_t st,
                         int nrhs = 1, int64_t vs = 0);

// word 4 of every matrix-core record <- s32[its element]
void launch_rec_scale(int nrec, const int *rec_elem, const float *s32, uint32_t *rec_words,
                      hipStream_t st);
// element stiffness scales for a level: s = penalty(rho) * 2^-level
void launch_simp_scale(int n, const double *rho, const int16_t *level, double q, double k0,
                       int penalty_type, double *s64, float *s32, hipStream_t st);
// coarse element scales: volume-weighted mean of the material factor of the
// fine elements inside each coarse element (CSR over coarse elements)
void launch_restrict_scale(int nc, const int *ptr, const int *child, const float *vw,
                           const double *s_fine, const int16_t *lev_fine,
                           const int16_t *lev_coarse, double *s64, float *s32, hipStream_t st);

// nodal design (TMROctConstitutive with one design variable per node,
// src/topology/TMROctConstitutive.cpp:498-551): rho at the 8 Gauss points of
// every element from its corner values (hanging corners interpolated from
// their parents), s = penalty(rho) * 2^-level per point; rho_gp (may be null)
// receives the point densities
void launch_nodal_scale(int nelem, const int *conn, const int *dep_ptr, const int *dep_conn,
                        const float *dep_w, const double *xn, const int16_t *level, double q,
                        double k0, int penalty_type, double *rho_gp, double *s64, float *s32,
                        hipStream_t st);
// design vector to the next coarser level: x_c = (R x_f) / (R 1), R = P^T
// (TACSBVecInterp::multWeightTranspose as used by TMRConformFilter::setDesignVars,
// src/topology/TMRConformFilter.cpp:154-167)
void launch_weight_restrict(const TransferCSR &R, const double *xf, double *xc, hipStream_t st);
// v_e^T K_g u_e at every Gauss point of every element (full-block tables)
void launch_nodal_energy(int nelem, const int *conn, const int *dep_ptr, const int *dep_conn,
                         const float *dep_w, const uint8_t *bcmask, const int *cls,
                         const double *G, const double *u, const double *v, double *out,
                         hipStream_t st);
// nodal design gradient into slot partials (element batches, component 0):
// t_a = sum_g N_a(g) (ex_g - dpenalty(rho_g) 2^-level e_g), ex may be null
void launch_nodal_grad(const ElemBatches &eb, const int16_t *lev, const double *rho_gp, double q,
                       int penalty_type, const double *e_gp, const double *ex_gp,
                       double *partial, hipStream_t st);
// out[i] = in[3 i]
void launch_take_comp0(int n, const double *in, double *out, hipStream_t st);

// FP64 PCG helpers (scalars live on the device)
// SC_XIT: the iteration whose alpha p is still to be added to x (-1: none)
enum { SC_RZ = 0, SC_PQ, SC_ALPHA, SC_RR, SC_ZQ, SC_BETA, SC_BB, SC_RZNEW, SC_TRR, SC_XIT, SC_N };
int reduce_blocks(int64_t n);
void launch_pcg_update_xr(int64_t ndof, double *x, double *r, const double *p, const double *q,
                          const double *sc, double *partials, float *r32, hipStream_t st);
void launch_dot2(int64_t ndof, const double *z, const double *r, const double *q,
                 double *partials, hipStream_t st);
void launch_dot1(int64_t ndof, const double *a, const double *b, double *partials,
                 hipStream_t st);
void launch_pcg_update_p(int64_t ndof, double *p, const double *z, const double *sc,
                         hipStream_t st);
void launch_dinv_dots(int64_t n, const double *dinv, double *y, const double *x,
                      double *partials, hipStream_t st);
// finalize: mode 0 rz, 1 pq->alpha, 2 rr, 3 (zr, zq)->beta, 4 bb,
// 5 (|y|^2, |x|^2) of the power iteration
void launch_finalize(int nparts, const double *partials, int mode, double *sc, hipStream_t st);

// Batched PCG over nrhs right-hand sides (vectors [nrhs][n]) whose dofs split
// into independent segments (components) with their own scalars
// sc[(rhs * nseg + seg) * SC_N + k], running flags act[rhs * nseg + seg] and
// fixed reduction blocks that never straddle a segment.
struct SegPlan {
  int nrhs = 1, nseg = 1, nblk = 0, ntiles = 0;
  int64_t n = 0;                 // dofs per right-hand side
  const int *blk_lo = nullptr;   // [nblk + 1] block b = dofs [blk_lo[b], blk_lo[b+1])
  const int *blk_seg = nullptr;  // [nblk]
  const int *seg_blk = nullptr;  // [nseg + 1]
  const int *seg_tile = nullptr; // [nseg + 1] finest-level owner tiles of each segment
};
enum { FIN_BB = 0, FIN_RZ, FIN_PQ, FIN_RR, FIN_BETA, FIN_TRUE };
void launch_b_dot1(const SegPlan &pl, const double *a, const double *b, const int *act,
                   double *partials, hipStream_t st);
// r -= alpha q, the FP32 copy r32 of r and, with x1, the finest level's
// first Chebyshev step from zero x1 = c0 Dinv32 r32 (x += alpha p follows in
// the same iteration's p update, which reads p anyway)
void launch_b_update_xr(const SegPlan &pl, double *r, const double *q, float *r32,
                        const double *sc, const int *act, double *partials, hipStream_t st,
                        const float *dinv32 = nullptr, double c0 = 0.0, float *x1 = nullptr);
template <class TZ>
void launch_b_dot2(const SegPlan &pl, const TZ *z, const double *r, const double *q,
                   const int *act, double *partials, hipStream_t st);
// x += alpha p for segments whose alpha was formed in the iteration before
// *iter (SC_XIT), then p = z + beta p for running segments
template <class TZ>
void launch_b_update_p(const SegPlan &pl, double *p, const TZ *z, const double *sc,
                       const int *act, double *x, const int *iter, hipStream_t st);
void launch_b_copy(const SegPlan &pl, const double *a, double *b, const int *act,
                   hipStream_t st);
// FIN_BB also sets the stop threshold max(rtol2 |b|^2, atol2) and the flags;
// FIN_RR stops a segment at the threshold (itconv = iteration, negative on
// breakdown); iter is the device iteration counter advanced by launch_b_tick
void launch_b_finalize(const SegPlan &pl, int mode, const double *partials, double *sc, int *act,
                       double *thr2, const int *iter, int *itconv, double rtol2, double atol2,
                       hipStream_t st);
// advance the iteration counter and count the running segments into *nact
void launch_b_tick(int *iter, const int *act, int nidx, int *nact, hipStream_t st);

// per-element compliance density and its design derivative (element mode)
void launch_elem_energy(int nelem, const int *conn, const int *dep_ptr, const int *dep_conn,
                        const float *dep_w, const uint8_t *bcmask, const int *cls,
                        const double *K0, const double *u, double *energy, hipStream_t st);
// p-norm of the relaxed von Mises failure (TMROctConstitutive::evalFailure,
// src/topology/TMROctConstitutive.cpp:1056-1104; TACS "pnorm-continuous"
// aggregation): S = sum_e sum_g (V_e / 8) (r(rho_e) vm(C0 B_g u_e) / ys)^P.
// CB: per class, the 8 Gauss-point stress operators C0 B_g (6 x 24) of the
// full-block element; vol: per class block volume.
struct StressArgs {
  double P = 8.0, eps = 0.3, ys = 1.0;
  const double *CB = nullptr;
  const double *vol = nullptr;
};
// rho_gp (nodal design, may be null): densities at the Gauss points, used
// in place of the element densities rho
void launch_stress_sum(int nelem, const int *conn, const int *dep_ptr, const int *dep_conn,
                       const float *dep_w, const uint8_t *bcmask, const int *cls,
                       const int16_t *lev, const double *rho, const double *rho_gp,
                       const double *u, const StressArgs &sa, double *partials, hipStream_t st);
// element dfdu into slot partials (element batches) and the explicit
// d f / d rho_e (or, with rho_gp, d f / d rho_g per Gauss point);
// coef = S^(1/P - 1) read from sc[SC_ZQ]
void launch_stress_grad(const ElemBatches &eb, const int16_t *lev, const int *cls,
                        const double *rho, const double *rho_gp, const double *u,
                        const StressArgs &sa, const double *sc, double *partial,
                        double *dfdrho, hipStream_t st);
// psi_e^T K0 u_e per element (adjoint-residual product without the factor)
void launch_elem_energy2(int nelem, const int *conn, const int *dep_ptr, const int *dep_conn,
                         const float *dep_w, const uint8_t *bcmask, const int *cls,
                         const double *K0, const double *u, const double *v, double *out,
                         hipStream_t st);
// g = dfdrho_explicit - dpenalty(rho) 2^-level * prod
void launch_adjoint_grad(int n, const double *rho, const int16_t *lev, double q, int ptype,
                         const double *prod, const double *explicit_part, double *g,
                         hipStream_t st);

// SIMP/RAMP derivative (src/topology/TMROctConstitutive.cpp:597-619) times
// -2^-level times the element energy
void launch_compliance_grad(int n, const double *rho, const int16_t *lev, double q, int ptype,
                            const double *energy, double *g, hipStream_t st);

}  // namespace tmrx
