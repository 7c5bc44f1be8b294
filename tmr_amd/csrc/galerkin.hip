// Galerkin coarse operators (TMR_CreateTACSMg with use_galerkin,
// src/TMR_RefinementTools.cpp:37-107: TACSMg forms P^T A P) on the GPU.
//
// Every element e of a finer level lies inside one element E of the next
// coarser level, and the prolongation restricted to e is the trilinear
// interpolation Q_e of E's corners at e's corners (hanging corners included:
// their constraint is the same trilinear interpolation on a coarser face or
// edge).  So P^T Z_f^T K_f Z_f P = Z_c^T (sum_E K_E) Z_c with the element
// matrices
//   K_E = sum_{e in E} (Q_e x I3)^T K_e (Q_e x I3),
// computed level by level (k_galerkin_elem), then assembled into a 3 x 3
// block-sparse (BSR) matrix over the independent nodes of the coarse level
// with the hanging-node expansion Z_c and identity rows and columns on the
// clamped dofs (k_bsr_assemble).  The block sums follow fixed term lists,
// so the operator is the same bit for bit on every setup.
#include <hip/hip_runtime.h>

#include "kernels.h"

namespace tmrx {

// ---------------------------------------------------------------------------
// K_E of every coarse element: one workgroup per element, the 576 entries
// spread over its threads.  MODE 0: children are finest-level elements of
// element densities, K_e = s_e K0[class]; MODE 1: children carry their own
// matrices Kin[e]; MODE 2: finest-level elements of a nodal design,
// K_e = sum_g s_eg Kg[class][g].
constexpr int GAL_THREADS = 256;

template <int MODE>
__global__ void __launch_bounds__(GAL_THREADS)
    k_galerkin_elem(int nE, const int *__restrict__ ptr, const int *__restrict__ child,
                    const int *__restrict__ place, const int *__restrict__ cls,
                    const double *__restrict__ s, const double *__restrict__ Ktab,
                    const double *__restrict__ Kin, double *__restrict__ Kout) {
  __shared__ double A[576], B[576], Q[64];
  const int E = blockIdx.x;
  if (E >= nE) return;
  const int tid = threadIdx.x;
  double acc[3] = {0.0, 0.0, 0.0};
  for (int k = ptr[E]; k < ptr[E + 1]; k++) {
    const int e = child[k];
    if (tid < 64) {
      // trilinear weight of coarse corner c at fine corner a
      const int pl = place[k];
      const int dl = (pl >> 24) & 15;
      const double inv = 1.0 / (double)(1 << dl);
      const int a = tid >> 3, c = tid & 7;
      const int off[3] = {pl & 255, (pl >> 8) & 255, (pl >> 16) & 255};
      double w = 1.0;
      for (int d = 0; d < 3; d++) {
        const double xi = (double)(off[d] + ((a >> d) & 1)) * inv;
        w *= ((c >> d) & 1) ? xi : 1.0 - xi;
      }
      Q[8 * a + c] = w;
    }
    for (int t = tid; t < 576; t += GAL_THREADS) {
      double v;
      if (MODE == 0) {
        v = s[e] * Ktab[576 * (size_t)cls[e] + t];
      } else if (MODE == 1) {
        v = Kin[576 * (size_t)e + t];
      } else {
        const double *Kg = Ktab + 576 * 8 * (size_t)cls[e];
        v = 0.0;
#pragma unroll
        for (int g = 0; g < 8; g++) v += s[8 * (size_t)e + g] * Kg[576 * g + t];
      }
      A[t] = v;
    }
    __syncthreads();
    // B = A (Q x I3): B[(3a+i)][(3c+j)] = sum_b A[(3a+i)][(3b+j)] Q[b][c]
    for (int t = tid; t < 576; t += GAL_THREADS) {
      const int row = t / 24, col = t % 24, c = col / 3, j = col % 3;
      double v = 0.0;
#pragma unroll
      for (int b = 0; b < 8; b++) v += A[24 * row + 3 * b + j] * Q[8 * b + c];
      B[t] = v;
    }
    __syncthreads();
    // K_E += (Q x I3)^T B
    int u = 0;
    for (int t = tid; t < 576; t += GAL_THREADS, u++) {
      const int row = t / 24, col = t % 24, c = row / 3, i = row % 3;
      double v = 0.0;
#pragma unroll
      for (int a = 0; a < 8; a++) v += Q[8 * a + c] * B[24 * (3 * a + i) + col];
      acc[u] += v;
    }
    __syncthreads();
  }
  int u = 0;
  for (int t = tid; t < 576; t += GAL_THREADS, u++) Kout[576 * (size_t)E + t] = acc[u];
}

void launch_galerkin_elem(int mode, int nE, const int *ptr, const int *child, const int *place,
                          const int *cls, const double *s, const double *Ktab, const double *Kin,
                          double *Kout, hipStream_t st) {
  if (nE == 0) return;
  if (mode == 0)
    hipLaunchKernelGGL(k_galerkin_elem<0>, dim3(nE), dim3(GAL_THREADS), 0, st, nE, ptr, child,
                       place, cls, s, Ktab, Kin, Kout);
  else if (mode == 1)
    hipLaunchKernelGGL(k_galerkin_elem<1>, dim3(nE), dim3(GAL_THREADS), 0, st, nE, ptr, child,
                       place, cls, s, Ktab, Kin, Kout);
  else
    hipLaunchKernelGGL(k_galerkin_elem<2>, dim3(nE), dim3(GAL_THREADS), 0, st, nE, ptr, child,
                       place, cls, s, Ktab, Kin, Kout);
  TMRX_LAUNCH_CHECK();
}

// element matrices from scales, one thread per (element, entry)
__global__ void k_elem_matrices(int ne, const int *__restrict__ cls, const double *__restrict__ s,
                                const double *__restrict__ Ktab, int nodal,
                                double *__restrict__ KE) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 576 * (int64_t)ne) return;
  const int e = (int)(t / 576), i = (int)(t % 576);
  double v;
  if (nodal) {
    const double *Kg = Ktab + 576 * 8 * (size_t)cls[e] + i;
    v = 0.0;
    for (int g = 0; g < 8; g++) v += s[8 * (size_t)e + g] * Kg[576 * g];
  } else {
    v = s[e] * Ktab[576 * (size_t)cls[e] + i];
  }
  KE[t] = v;
}

void launch_elem_matrices(int ne, const int *cls, const double *s, const double *Ktab, bool nodal,
                          double *KE, hipStream_t st) {
  if (ne == 0) return;
  const int64_t n = 576 * (int64_t)ne;
  hipLaunchKernelGGL(k_elem_matrices, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, ne, cls,
                     s, Ktab, nodal ? 1 : 0, KE);
  TMRX_LAUNCH_CHECK();
}

// ---------------------------------------------------------------------------
// BSR values: block k (row node i = brow[k], column node j = col[k]) is the
// sum over its terms of w * K_E[(3a+p)][(3b+q)], term code = 64 E + 8 a + b;
// rows and columns of clamped dofs become identity rows and columns.
__global__ void k_bsr_assemble(int nblk, const int *__restrict__ tptr, const int *__restrict__ tcode,
                               const double *__restrict__ tw, const int *__restrict__ brow,
                               const int *__restrict__ col, const double *__restrict__ KE,
                               const uint8_t *__restrict__ bcmask, double *__restrict__ val) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nblk) return;
  double v[9];
#pragma unroll
  for (int q = 0; q < 9; q++) v[q] = 0.0;
  for (int t = tptr[k]; t < tptr[k + 1]; t++) {
    const int code = tcode[t];
    const double w = tw[t];
    const double *K = KE + 576 * (size_t)(code >> 6) + 24 * (3 * ((code >> 3) & 7)) +
                      3 * (code & 7);
#pragma unroll
    for (int p = 0; p < 3; p++)
#pragma unroll
      for (int q = 0; q < 3; q++) v[3 * p + q] += w * K[24 * p + q];
  }
  const int i = brow[k], j = col[k];
  const uint8_t mi = bcmask[i], mj = bcmask[j];
#pragma unroll
  for (int p = 0; p < 3; p++)
#pragma unroll
    for (int q = 0; q < 3; q++)
      if (((mi >> p) & 1) || ((mj >> q) & 1)) v[3 * p + q] = (i == j && p == q) ? 1.0 : 0.0;
#pragma unroll
  for (int q = 0; q < 9; q++) val[9 * (size_t)k + q] = v[q];
}

void launch_bsr_assemble(const BSR64 &A, const int *tptr, const int *tcode, const double *tw,
                         const int *brow, const double *KE, double *val, hipStream_t st) {
  if (A.nnzb == 0) return;
  hipLaunchKernelGGL(k_bsr_assemble, dim3((A.nnzb + 255) / 256), dim3(256), 0, st, A.nnzb, tptr,
                     tcode, tw, brow, A.col, KE, A.bcmask, val);
  TMRX_LAUNCH_CHECK();
}

// Jacobi diagonal from the diagonal blocks (clamped dofs have 1 there)
template <class T>
__global__ void k_bsr_dinv(int n, const int *__restrict__ diag, const double *__restrict__ val,
                           T *__restrict__ dinv) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double *v = val + 9 * (size_t)diag[i];
#pragma unroll
  for (int p = 0; p < 3; p++) dinv[3 * (size_t)i + p] = (T)(1.0 / v[4 * p]);
}

template <class T>
void launch_bsr_dinv(int n, const int *diag, const double *val, T *dinv, hipStream_t st) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_bsr_dinv<T>, dim3((n + 255) / 256), dim3(256), 0, st, n, diag, val, dinv);
  TMRX_LAUNCH_CHECK();
}

// ---------------------------------------------------------------------------
// BSR operator with the epilogues of the tile apply (ASM_STORE, ASM_RESID,
// ASM_CHEB): BSR_LANES lanes per block row, each summing every BSR_LANES-th
// block of the row, then a butterfly over the group in a fixed order.
constexpr int BSR_LANES = 4;

template <class T, int MODE>
__global__ void __launch_bounds__(256)
    k_bsr_apply(int n, const int *__restrict__ ptr, const int *__restrict__ col,
                const T *__restrict__ val, const uint8_t *__restrict__ bcmask,
                const T *__restrict__ x, const T *__restrict__ b, T *__restrict__ y,
                T *__restrict__ x_out, T *__restrict__ d, const T *__restrict__ dinv, T ca, T cc,
                int cflags, const int *__restrict__ run, int64_t vs) {
  if (run && *run == 0) return;
  const int64_t off = (int64_t)blockIdx.y * vs;
  x += off;
  if (b) b += off;
  if (y) y += off;
  if (x_out) x_out += off;
  if (d) d += off;
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  const int row = g / BSR_LANES, sub = g % BSR_LANES;
  const bool valid = row < n;
  T a0 = T(0), a1 = T(0), a2 = T(0);
  if (valid) {
    const int k1 = ptr[row + 1];
    for (int k = ptr[row] + sub; k < k1; k += BSR_LANES) {
      const int j = col[k];
      const T *v = val + 9 * (size_t)k;
      const T x0 = x[3 * (size_t)j], x1 = x[3 * (size_t)j + 1], x2 = x[3 * (size_t)j + 2];
      a0 += v[0] * x0 + v[1] * x1 + v[2] * x2;
      a1 += v[3] * x0 + v[4] * x1 + v[5] * x2;
      a2 += v[6] * x0 + v[7] * x1 + v[8] * x2;
    }
  }
#pragma unroll
  for (int m = 1; m < BSR_LANES; m <<= 1) {
    a0 += __shfl_xor(a0, m);
    a1 += __shfl_xor(a1, m);
    a2 += __shfl_xor(a2, m);
  }
  if (!valid || sub != 0) return;
  const size_t k3 = 3 * (size_t)row;
  const uint8_t msk = bcmask[row];
  const T a[3] = {a0, a1, a2};
  T yo[3], dno[3];
#pragma unroll
  for (int c = 0; c < 3; c++) {
    const bool bc = (msk >> c) & 1;
    if (MODE == ASM_STORE) {
      yo[c] = bc ? x[k3 + c] : a[c];
    } else if (MODE == ASM_RESID) {
      yo[c] = bc ? T(0) : b[k3 + c] - a[c];
    } else {  // ASM_CHEB
      if (bc) {
        dno[c] = T(0);
        yo[c] = T(0);
      } else {
        const T r = b[k3 + c] - a[c];
        const T xc = x[k3 + c];
        const T dp = (cflags & CHEB_D_IS_X) ? xc : ((cflags & CHEB_D_READ) ? d[k3 + c] : T(0));
        const T dn = ca * dp + cc * dinv[k3 + c] * r;
        dno[c] = dn;
        yo[c] = xc + dn;
      }
    }
  }
  if (MODE == ASM_CHEB) {
    if (cflags & CHEB_D_STORE)
#pragma unroll
      for (int c = 0; c < 3; c++) d[k3 + c] = dno[c];
#pragma unroll
    for (int c = 0; c < 3; c++) x_out[k3 + c] = yo[c];
  } else {
#pragma unroll
    for (int c = 0; c < 3; c++) y[k3 + c] = yo[c];
  }
}

template <class T>
void launch_bsr_apply(const BSR<T> &A, const T *x, int mode, const AsmArgs &g, hipStream_t st,
                      int nrhs, int64_t vs) {
  if (A.n == 0) return;
  const dim3 grid((unsigned)(((int64_t)A.n * BSR_LANES + 255) / 256), (unsigned)nrhs);
  const T *b = (const T *)g.b;
  T *y = (T *)g.y, *xo = (T *)g.x_out, *d = (T *)g.d;
  const T *dinv = (const T *)g.dinv;
  const T ca = (T)g.ca, cc = (T)g.cc;
  if (mode == ASM_STORE)
    hipLaunchKernelGGL((k_bsr_apply<T, ASM_STORE>), grid, dim3(256), 0, st, A.n, A.ptr, A.col,
                       A.val, A.bcmask, x, b, y, xo, d, dinv, ca, cc, g.cflags, A.run, vs);
  else if (mode == ASM_RESID)
    hipLaunchKernelGGL((k_bsr_apply<T, ASM_RESID>), grid, dim3(256), 0, st, A.n, A.ptr, A.col,
                       A.val, A.bcmask, x, b, y, xo, d, dinv, ca, cc, g.cflags, A.run, vs);
  else if (mode == ASM_CHEB)
    hipLaunchKernelGGL((k_bsr_apply<T, ASM_CHEB>), grid, dim3(256), 0, st, A.n, A.ptr, A.col,
                       A.val, A.bcmask, x, b, y, xo, d, dinv, ca, cc, g.cflags, A.run, vs);
  else
    note_launch(hipErrorInvalidValue);
  TMRX_LAUNCH_CHECK();
}

// ---------------------------------------------------------------------------
// One colour of a point SOR sweep on the BSR matrix (TACSGaussSeidel's
// smoother, symmetric SOR with relaxation omega; TMR_CreateTACSMg,
// src/TMR_RefinementTools.cpp:93-94).  The matrix is the level's BSR with
// its rows in the sweep's colour-major node order (row k holds node
// nodes[k]; columns are node indices), so a colour's rows are contiguous.
// Every node of the colour takes
// SOR_LANES lanes: they sum the row's off-diagonal blocks (consecutive
// lanes on consecutive blocks, which read x of other colours only, so the
// nodes of a colour update side by side) and a butterfly combines them;
// then the node's three dofs update one after another through the diagonal
// block, 0, 1, 2 (BACK: 2, 1, 0), each
// x_p <- (1 - omega) x_p + omega (b_p - sum_{q != p} a_pq x_q) / a_pp --
// the point SOR of the dofs in colour-major node order.
constexpr int SOR_LANES = 8;

template <class T, bool BACK>
__global__ __launch_bounds__(256) void k_sor_colour(int cnt, const int *__restrict__ nodes,
                                                    const int *__restrict__ ptr,
                                                    const int *__restrict__ col,
                                                    const T *__restrict__ val,
                                                    const T *__restrict__ b, T *__restrict__ x,
                                                    T omega, const int *__restrict__ run,
                                                    int64_t vs) {
  if (run && *run == 0) return;
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  const int k = g / SOR_LANES, sub = g % SOR_LANES;
  const bool valid = k < cnt;
  const int64_t off = (int64_t)blockIdx.y * vs;
  x += off;
  b += off;
  const int i = valid ? nodes[k] : 0;
  T s0 = T(0), s1 = T(0), s2 = T(0);
  int kd = -1;
  if (valid) {
    const int q1 = ptr[k + 1];
    for (int q = ptr[k] + sub; q < q1; q += SOR_LANES) {
      const int j = col[q];
      if (j == i) {
        kd = q;
        continue;
      }
      const T *v = val + 9 * (size_t)q;
      const T x0 = x[3 * (size_t)j], x1 = x[3 * (size_t)j + 1], x2 = x[3 * (size_t)j + 2];
      s0 += v[0] * x0 + v[1] * x1 + v[2] * x2;
      s1 += v[3] * x0 + v[4] * x1 + v[5] * x2;
      s2 += v[6] * x0 + v[7] * x1 + v[8] * x2;
    }
  }
#pragma unroll
  for (int m = 1; m < SOR_LANES; m <<= 1) {
    s0 += __shfl_xor(s0, m);
    s1 += __shfl_xor(s1, m);
    s2 += __shfl_xor(s2, m);
    kd = max(kd, __shfl_xor(kd, m));
  }
  if (!valid || sub != 0 || kd < 0) return;
  const T *D = val + 9 * (size_t)kd;
  const size_t i3 = 3 * (size_t)i;
  T xi[3] = {x[i3], x[i3 + 1], x[i3 + 2]};
  const T r[3] = {b[i3] - s0, b[i3 + 1] - s1, b[i3 + 2] - s2};
#pragma unroll
  for (int t = 0; t < 3; t++) {
    const int p = BACK ? 2 - t : t;
    T acc = r[p];
#pragma unroll
    for (int q = 0; q < 3; q++)
      if (q != p) acc -= D[3 * p + q] * xi[q];
    xi[p] = (T(1) - omega) * xi[p] + omega * acc / D[4 * p];
  }
#pragma unroll
  for (int c = 0; c < 3; c++) x[i3 + c] = xi[c];
}

// val_out[k] = val[map[k]] for every 3 x 3 block (the colour-major copy)
__global__ void k_bsr_gather(int nnzb, const int *__restrict__ map, const double *__restrict__ val,
                             double *__restrict__ out) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 9 * (int64_t)nnzb) return;
  const int64_t k = t / 9, e = t % 9;
  out[t] = val[9 * (int64_t)map[k] + e];
}

void launch_bsr_gather(int nnzb, const int *map, const double *val, double *out, hipStream_t st) {
  if (nnzb == 0) return;
  const int64_t n = 9 * (int64_t)nnzb;
  hipLaunchKernelGGL(k_bsr_gather, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, nnzb, map,
                     val, out);
  TMRX_LAUNCH_CHECK();
}

template <class T>
void launch_sor_sweep(const BSR<T> &A, const int *nodes, const int *cptr, int ncolors, bool back,
                      const T *b, T *x, double omega, hipStream_t st, int nrhs, int64_t vs) {
  for (int t = 0; t < ncolors; t++) {
    const int c = back ? ncolors - 1 - t : t;
    const int c0 = cptr[c], cnt = cptr[c + 1] - c0;
    if (cnt == 0) continue;
    const dim3 grid((unsigned)(((int64_t)cnt * SOR_LANES + 255) / 256), (unsigned)nrhs);
    if (back)
      hipLaunchKernelGGL((k_sor_colour<T, true>), grid, dim3(256), 0, st, cnt, nodes + c0,
                         A.ptr + c0, A.col, A.val, b, x, (T)omega, A.run, vs);
    else
      hipLaunchKernelGGL((k_sor_colour<T, false>), grid, dim3(256), 0, st, cnt, nodes + c0,
                         A.ptr + c0, A.col, A.val, b, x, (T)omega, A.run, vs);
    TMRX_LAUNCH_CHECK();
  }
}

template void launch_sor_sweep<float>(const BSR<float> &, const int *, const int *, int, bool,
                                      const float *, float *, double, hipStream_t, int, int64_t);
template void launch_sor_sweep<double>(const BSR<double> &, const int *, const int *, int, bool,
                                       const double *, double *, double, hipStream_t, int,
                                       int64_t);
template void launch_bsr_dinv<float>(int, const int *, const double *, float *, hipStream_t);
template void launch_bsr_dinv<double>(int, const int *, const double *, double *, hipStream_t);
template void launch_bsr_apply<float>(const BSR<float> &, const float *, int, const AsmArgs &,
                                      hipStream_t, int, int64_t);
template void launch_bsr_apply<double>(const BSR<double> &, const double *, int, const AsmArgs &,
                                       hipStream_t, int, int64_t);

}  // namespace tmrx
