// HIP kernels for gfx950 (wave64).  See kernels.h for the data layout.
#include <hip/hip_runtime.h>

#include "kernels.h"

namespace tmrx {

// the first failed launch is kept until a synchronising call collects it
static hipError_t g_launch_error = hipSuccess;
void note_launch(hipError_t e) {
  if (e != hipSuccess && g_launch_error == hipSuccess) g_launch_error = e;
}
hipError_t take_launch_error() {
  const hipError_t e = g_launch_error;
  g_launch_error = hipSuccess;
  return e;
}

// Blocks are dealt to the 8 XCDs round-robin; remap so that each XCD walks a
// contiguous range of logical blocks (neighbouring Morton batches share nodes,
// which then hit the same L2).
__device__ __forceinline__ int xcd_block(int b, int n) {
  const int x = b & 7, i = b >> 3, q = n >> 3, r = n & 7;
  return x * q + (x < r ? x : r) + i;
}

// ---------------------------------------------------------------------------
// element pass: stage batch nodes in LDS, y_e = s_e K0 x_e, colored LDS
// accumulation, one partial per slot.
// the three components of one node as a single 12-byte (FP32) or
// 24-byte (FP64) load
template <class T>
struct __attribute__((packed, aligned(sizeof(T)))) Vec3 {
  T a, b, c;
};
template <class T>
__device__ __forceinline__ Vec3<T> load_vec3(const T *__restrict__ p) {
  return *reinterpret_cast<const Vec3<T> *>(p);
}
template <class T>
__device__ __forceinline__ void store_vec3(T *__restrict__ p, const T v[3]) {
  *reinterpret_cast<Vec3<T> *>(p) = Vec3<T>{v[0], v[1], v[2]};
}

template <class T>
__device__ __forceinline__ void load_node(int e, const T *__restrict__ x,
                                          const int *__restrict__ dep_ptr,
                                          const int *__restrict__ dep_conn,
                                          const float *__restrict__ dep_w,
                                          const uint8_t *__restrict__ bcmask, T v[3]) {
  if (e >= 0) {
    const uint8_t m = bcmask[e];
    const Vec3<T> u = load_vec3(x + 3 * (size_t)e);
    v[0] = (m & 1) ? T(0) : u.a;
    v[1] = (m & 2) ? T(0) : u.b;
    v[2] = (m & 4) ? T(0) : u.c;
  } else {
    const int d = -e - 1;
    v[0] = v[1] = v[2] = T(0);
    for (int j = dep_ptr[d]; j < dep_ptr[d + 1]; j++) {
      const int p = dep_conn[j];
      const T w = (T)dep_w[j];
      const uint8_t m = bcmask[p];
      const Vec3<T> u = load_vec3(x + 3 * (size_t)p);
      if (!(m & 1)) v[0] += w * u.a;
      if (!(m & 2)) v[1] += w * u.b;
      if (!(m & 4)) v[2] += w * u.c;
    }
  }
}

template <class T>
__global__ __launch_bounds__(WG_ELEMS) void k_elem_apply(
    const int *__restrict__ wg_node_ptr, const int *__restrict__ wg_nodes,
    const int *__restrict__ wg_elem_ptr, const int *__restrict__ wg_class,
    const uint4 *__restrict__ lconn, const uint8_t *__restrict__ color, int ncolors,
    const int *__restrict__ dep_ptr, const int *__restrict__ dep_conn,
    const float *__restrict__ dep_w, const uint8_t *__restrict__ bcmask,
    const T *__restrict__ K0, const T *__restrict__ s_elem, const T *__restrict__ x,
    T *__restrict__ partial) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int wg = xcd_block(blockIdx.x, gridDim.x);
  const int tid = threadIdx.x;
  const int nb = wg_node_ptr[wg];
  const int nn = wg_node_ptr[wg + 1] - nb;
  T *xs = reinterpret_cast<T *>(smem_raw);
  T *ys = xs + 3 * nn;
  for (int i = tid; i < nn; i += WG_ELEMS) {
    T v[3];
    load_node<T>(wg_nodes[nb + i], x, dep_ptr, dep_conn, dep_w, bcmask, v);
    xs[3 * i] = v[0];
    xs[3 * i + 1] = v[1];
    xs[3 * i + 2] = v[2];
    ys[3 * i] = ys[3 * i + 1] = ys[3 * i + 2] = T(0);
  }
  __syncthreads();
  const int e0 = wg_elem_ptr[wg];
  const int ne = wg_elem_ptr[wg + 1] - e0;
  const int cls = __builtin_amdgcn_readfirstlane(wg_class[wg]);
  const T *__restrict__ K = K0 + 576 * cls;
  T ye[24];
  int lc[8];
  int col = -1;
  if (tid < ne) {
    const uint4 q = lconn[e0 + tid];
    lc[0] = q.x & 0xffff; lc[1] = q.x >> 16;
    lc[2] = q.y & 0xffff; lc[3] = q.y >> 16;
    lc[4] = q.z & 0xffff; lc[5] = q.z >> 16;
    lc[6] = q.w & 0xffff; lc[7] = q.w >> 16;
    col = color[e0 + tid];
    const T s = s_elem[e0 + tid];
    T xe[24];
#pragma unroll
    for (int a = 0; a < 8; a++) {
      xe[3 * a] = xs[3 * lc[a]];
      xe[3 * a + 1] = xs[3 * lc[a] + 1];
      xe[3 * a + 2] = xs[3 * lc[a] + 2];
    }
#pragma unroll
    for (int i = 0; i < 24; i++) {
      T acc = T(0);
#pragma unroll
      for (int j = 0; j < 24; j++) acc += K[24 * i + j] * xe[j];
      ye[i] = s * acc;
    }
  }
  for (int c = 0; c < ncolors; c++) {
    if (col == c) {
#pragma unroll
      for (int a = 0; a < 8; a++) {
        ys[3 * lc[a]] += ye[3 * a];
        ys[3 * lc[a] + 1] += ye[3 * a + 1];
        ys[3 * lc[a] + 2] += ye[3 * a + 2];
      }
    }
    __syncthreads();
  }
  for (int i = tid; i < 3 * nn; i += WG_ELEMS) partial[3 * nb + i] = ys[i];
}

template <class T, bool NODAL>
__global__ __launch_bounds__(WG_ELEMS) void k_elem_diag(
    const int *__restrict__ wg_node_ptr, const int *__restrict__ wg_elem_ptr,
    const int *__restrict__ wg_class, const uint4 *__restrict__ lconn,
    const uint8_t *__restrict__ color, int ncolors, const T *__restrict__ K0,
    const T *__restrict__ s_elem, T *__restrict__ partial) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int wg = blockIdx.x, tid = threadIdx.x;
  const int nb = wg_node_ptr[wg], nn = wg_node_ptr[wg + 1] - nb;
  T *ys = reinterpret_cast<T *>(smem_raw);
  for (int i = tid; i < 3 * nn; i += WG_ELEMS) ys[i] = T(0);
  __syncthreads();
  const int e0 = wg_elem_ptr[wg], ne = wg_elem_ptr[wg + 1] - e0;
  const int cls = __builtin_amdgcn_readfirstlane(wg_class[wg]);
  int lc[8];
  int col = -1;
  T de[24];
  if (tid < ne) {
    const uint4 q = lconn[e0 + tid];
    lc[0] = q.x & 0xffff; lc[1] = q.x >> 16;
    lc[2] = q.y & 0xffff; lc[3] = q.y >> 16;
    lc[4] = q.z & 0xffff; lc[5] = q.z >> 16;
    lc[6] = q.w & 0xffff; lc[7] = q.w >> 16;
    col = color[e0 + tid];
    if constexpr (NODAL) {
      // diag K_g = |J| w [(lambda + mu) G_d^2 + mu |G|^2] per Gauss point
      const T *__restrict__ Gt = K0 + NODAL_TABLE * cls;
      const T lam = Gt[192], mu = Gt[193];
      const T *__restrict__ sg = s_elem + 8 * (size_t)(e0 + tid);
#pragma unroll
      for (int i = 0; i < 24; i++) de[i] = T(0);
#pragma unroll
      for (int g = 0; g < 8; g++) {
        const T s = sg[g];
#pragma unroll
        for (int a = 0; a < 8; a++) {
          const T g0 = Gt[24 * g + 3 * a], g1 = Gt[24 * g + 3 * a + 1], g2 = Gt[24 * g + 3 * a + 2];
          const T n2 = mu * (g0 * g0 + g1 * g1 + g2 * g2);
          de[3 * a] += s * ((lam + mu) * g0 * g0 + n2);
          de[3 * a + 1] += s * ((lam + mu) * g1 * g1 + n2);
          de[3 * a + 2] += s * ((lam + mu) * g2 * g2 + n2);
        }
      }
    } else {
      const T *__restrict__ K = K0 + 576 * cls;
      const T s = s_elem[e0 + tid];
#pragma unroll
      for (int i = 0; i < 24; i++) de[i] = s * K[i * 25];
    }
  }
  for (int c = 0; c < ncolors; c++) {
    if (col == c) {
#pragma unroll
      for (int a = 0; a < 8; a++)
#pragma unroll
        for (int d = 0; d < 3; d++) ys[3 * lc[a] + d] += de[3 * a + d];
    }
    __syncthreads();
  }
  for (int i = tid; i < 3 * nn; i += WG_ELEMS) partial[3 * nb + i] = ys[i];
}

template <class T>
void launch_elem_apply(const ElemBatches &eb, const T *K0, const T *s_elem, const T *x,
                       T *partial, hipStream_t st) {
  if (eb.n_wg == 0) return;
  size_t sh = 2 * 3 * (size_t)eb.max_wg_nodes * sizeof(T);
  hipLaunchKernelGGL(k_elem_apply<T>, dim3(eb.n_wg), dim3(WG_ELEMS), sh, st, eb.wg_node_ptr,
                     eb.wg_nodes, eb.wg_elem_ptr, eb.wg_class,
                     reinterpret_cast<const uint4 *>(eb.lconn), eb.color, eb.ncolors,
                     eb.dep_ptr, eb.dep_conn, eb.dep_w, eb.bcmask, K0, s_elem, x, partial);
  TMRX_LAUNCH_CHECK();
}

template <class T>
void launch_elem_diag(const ElemBatches &eb, const T *K0, const T *s_elem, T *partial,
                      hipStream_t st, bool nodal) {
  if (eb.n_wg == 0) return;
  size_t sh = 3 * (size_t)eb.max_wg_nodes * sizeof(T);
  if (nodal)
    hipLaunchKernelGGL((k_elem_diag<T, true>), dim3(eb.n_wg), dim3(WG_ELEMS), sh, st,
                       eb.wg_node_ptr, eb.wg_elem_ptr, eb.wg_class,
                       reinterpret_cast<const uint4 *>(eb.lconn), eb.color, eb.ncolors, K0, s_elem,
                       partial);
  else
    hipLaunchKernelGGL((k_elem_diag<T, false>), dim3(eb.n_wg), dim3(WG_ELEMS), sh, st,
                       eb.wg_node_ptr, eb.wg_elem_ptr, eb.wg_class,
                       reinterpret_cast<const uint4 *>(eb.lconn), eb.color, eb.ncolors, K0, s_elem,
                       partial);
  TMRX_LAUNCH_CHECK();
}

// corner-pair diagonal terms (elements with hanging corners), one thread
// per node that has them; records in a fixed order, so the sums are too
template <bool NODAL>
__global__ __launch_bounds__(256) void k_diag_cross(
    int n, const int *__restrict__ ptr, const int *__restrict__ node, const int *__restrict__ code,
    const double *__restrict__ w, const int *__restrict__ cls, const double *__restrict__ K0,
    const double *__restrict__ s, const uint8_t *__restrict__ bcmask, double *__restrict__ dinv) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double c[3] = {0.0, 0.0, 0.0};
  for (int j = ptr[i]; j < ptr[i + 1]; j++) {
    const int cd = code[j], e = cd >> 6, a = (cd >> 3) & 7, b = cd & 7;
    const double wj = w[j];
    if (NODAL) {
      // K_g[3a+d][3b+d] = |J| w [(lambda + mu) G_ad G_bd + mu G_a . G_b]
      const double *__restrict__ Gt = K0 + NODAL_TABLE * cls[e];
      const double lam = Gt[192], mu = Gt[193];
      for (int g = 0; g < 8; g++) {
        const double *ga = Gt + 24 * g + 3 * a, *gb = Gt + 24 * g + 3 * b;
        const double sg = wj * s[8 * (size_t)e + g];
        const double dt = mu * (ga[0] * gb[0] + ga[1] * gb[1] + ga[2] * gb[2]);
        for (int d = 0; d < 3; d++) c[d] += sg * ((lam + mu) * ga[d] * gb[d] + dt);
      }
    } else {
      const double *__restrict__ K = K0 + 576 * cls[e];
      const double se = wj * s[e];
      for (int d = 0; d < 3; d++) c[d] += se * K[(3 * a + d) * 24 + 3 * b + d];
    }
  }
  const int nd = node[i];
  const uint8_t m = bcmask[nd];
  for (int d = 0; d < 3; d++)
    if (!((m >> d) & 1)) {
      const size_t k = 3 * (size_t)nd + d;
      dinv[k] = 1.0 / (1.0 / dinv[k] + c[d]);
    }
}

void launch_diag_cross(int n, const int *ptr, const int *node, const int *code, const double *w,
                       const int *cls, const double *K0, const double *s, bool nodal,
                       const uint8_t *bcmask, double *dinv, hipStream_t st) {
  if (n == 0) return;
  if (nodal)
    hipLaunchKernelGGL(k_diag_cross<true>, dim3((n + 255) / 256), dim3(256), 0, st, n, ptr, node,
                       code, w, cls, K0, s, bcmask, dinv);
  else
    hipLaunchKernelGGL(k_diag_cross<false>, dim3((n + 255) / 256), dim3(256), 0, st, n, ptr, node,
                       code, w, cls, K0, s, bcmask, dinv);
  TMRX_LAUNCH_CHECK();
}

// ---------------------------------------------------------------------------
// assembly pass: node-parallel weighted gather of slot partials + epilogue
constexpr int NODE_BLOCK = 256;

__device__ __forceinline__ double block_sum(double v, double *red) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) red[wave] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0) {
    for (int w = 0; w < (int)(blockDim.x >> 6); w++) s += red[w];
  }
  return s;
}

template <class T, int MODE>
__global__ __launch_bounds__(NODE_BLOCK) void k_assemble(
    int n, const int *__restrict__ ptr, const int *__restrict__ slot, const float *__restrict__ w,
    const uint8_t *__restrict__ bcmask, const T *__restrict__ partial, const T *__restrict__ b,
    T *__restrict__ y, T *__restrict__ x, T *__restrict__ d, const T *__restrict__ dinv,
    double ca, double cc, double *__restrict__ partials) {
  __shared__ double red[NODE_BLOCK / 64];
  const int i = xcd_block(blockIdx.x, gridDim.x) * NODE_BLOCK + threadIdx.x;
  double dot = 0.0;
  if (i < n) {
    T a0 = T(0), a1 = T(0), a2 = T(0);
    for (int j = ptr[i]; j < ptr[i + 1]; j++) {
      const int s = slot[j];
      T ww = (T)w[j];
      if (MODE == ASM_DIAG) ww = ww * ww;
      a0 += ww * partial[3 * s];
      a1 += ww * partial[3 * s + 1];
      a2 += ww * partial[3 * s + 2];
    }
    const uint8_t m = bcmask[i];
    const T acc[3] = {a0, a1, a2};
#pragma unroll
    for (int c = 0; c < 3; c++) {
      const int k = 3 * i + c;
      const bool bc = (m >> c) & 1;
      if (MODE == ASM_STORE) {
        y[k] = bc ? x[k] : acc[c];
      } else if (MODE == ASM_DOT) {
        const T v = bc ? x[k] : acc[c];
        y[k] = v;
        dot += (double)x[k] * (double)v;
      } else if (MODE == ASM_RESID) {
        y[k] = bc ? T(0) : b[k] - acc[c];
      } else if (MODE == ASM_CHEB) {
        if (bc) {
          d[k] = T(0);
          x[k] = T(0);
        } else {
          const T r = b[k] - acc[c];
          const T dn = (T)ca * d[k] + (T)cc * dinv[k] * r;
          d[k] = dn;
          x[k] += dn;
        }
      } else if (MODE == ASM_DIAG) {
        y[k] = bc ? T(1) : T(1) / acc[c];
      } else if (MODE == ASM_GATHER) {
        y[k] = bc ? T(0) : acc[c];
      }
    }
  }
  if (MODE == ASM_DOT) {
    const double s = block_sum(dot, red);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
  }
}

int assemble_num_blocks(int n_nodes) { return (n_nodes + NODE_BLOCK - 1) / NODE_BLOCK; }

template <class T>
void launch_assemble(const AsmCSR &a, const uint8_t *bcmask, const T *partial, int mode,
                     const AsmArgs &g, hipStream_t st) {
  const int nb = assemble_num_blocks(a.n_nodes);
  if (nb == 0) return;
  const T *b = (const T *)g.b;
  T *y = (T *)g.y, *x = (T *)g.x, *d = (T *)g.d;
  const T *dinv = (const T *)g.dinv;
#define TMRX_ASM(M)                                                                           \
  hipLaunchKernelGGL((k_assemble<T, M>), dim3(nb), dim3(NODE_BLOCK), 0, st, a.n_nodes, a.ptr, \
                     a.slot, a.w, bcmask, partial, b, y, x, d, dinv, g.ca, g.cc, g.partials)
  switch (mode) {
    case ASM_STORE: TMRX_ASM(ASM_STORE); break;
    case ASM_RESID: TMRX_ASM(ASM_RESID); break;
    case ASM_CHEB: TMRX_ASM(ASM_CHEB); break;
    case ASM_DIAG: TMRX_ASM(ASM_DIAG); break;
    case ASM_DOT: TMRX_ASM(ASM_DOT); break;
    case ASM_GATHER: TMRX_ASM(ASM_GATHER); break;
  }
#undef TMRX_ASM
  TMRX_LAUNCH_CHECK();
}

// ---------------------------------------------------------------------------
// grid transfer: prolongation (add into fine), restriction (coarse = P^T r)
template <class T>
__global__ __launch_bounds__(256) void k_prolong_add(int n, const int *__restrict__ ptr,
                                                     const int *__restrict__ col,
                                                     const float *__restrict__ w,
                                                     const uint8_t *__restrict__ mask,
                                                     const T *__restrict__ xc, T *__restrict__ xf,
                                                     int64_t vsc, int64_t vsf,
                                                     const int *__restrict__ run,
                                                     const uint8_t *__restrict__ wexp) {
  if (run && *run == 0) return;
  const int i = xcd_block(blockIdx.x, gridDim.x) * 256 + threadIdx.x;
  if (i >= n) return;
  xc += (int64_t)blockIdx.y * vsc;
  xf += (int64_t)blockIdx.y * vsf;
  const int e = wexp ? (int)wexp[i] : 0;
  const T wu = e ? (T)ldexpf(1.0f, 1 - e) : T(0);  // equal weights 2^-(e-1)
  T a0 = T(0), a1 = T(0), a2 = T(0);
  const Vec3<T> f = load_vec3(xf + 3 * (size_t)i);
  const uint8_t m = mask[i];
  const int j0 = ptr[i], j1 = ptr[i + 1];
  // four entries at a time: their index and value loads go out together
  // (indices clamped to the row's last entry; the extra terms are skipped)
  for (int j = j0; j < j1; j += 4) {
    int c[4];
#pragma unroll
    for (int u = 0; u < 4; u++) c[u] = col[min(j + u, j1 - 1)];
    Vec3<T> v[4];
#pragma unroll
    for (int u = 0; u < 4; u++) v[u] = load_vec3(xc + 3 * (size_t)c[u]);
#pragma unroll
    for (int u = 0; u < 4; u++)
      if (j + u < j1) {
        const T ww = e ? wu : (T)w[j + u];
        a0 += ww * v[u].a;
        a1 += ww * v[u].b;
        a2 += ww * v[u].c;
      }
  }
  const T o[3] = {(m & 1) ? f.a : f.a + a0, (m & 2) ? f.b : f.b + a1, (m & 4) ? f.c : f.c + a2};
  store_vec3(xf + 3 * (size_t)i, o);
}

// 8 lanes per coarse row: rows gather ~27 scattered fine nodes, so spread
// the loads over lanes and reduce with shuffles (fixed order).
template <class T>
__global__ __launch_bounds__(256) void k_restrict(int n, const int *__restrict__ ptr,
                                                  const int *__restrict__ col,
                                                  const float *__restrict__ w,
                                                  const uint8_t *__restrict__ mask,
                                                  const T *__restrict__ rf, T *__restrict__ rc,
                                                  int64_t vsf, int64_t vsc,
                                                  const int *__restrict__ run,
                                                  const T *__restrict__ dinv_c, double c0,
                                                  T *__restrict__ x1,
                                                  const uint8_t *__restrict__ wexp) {
  if (run && *run == 0) return;
  const int g = xcd_block(blockIdx.x, gridDim.x) * 256 + threadIdx.x;
  rf += (int64_t)blockIdx.y * vsf;
  rc += (int64_t)blockIdx.y * vsc;
  if (x1) x1 += (int64_t)blockIdx.y * vsc;
  const int i = g >> 3, sub = g & 7;
  T a0 = T(0), a1 = T(0), a2 = T(0);
  // the row's own inputs (mask, coarse diagonal) load before the gathers
  uint8_t m = 0;
  Vec3<T> di{T(0), T(0), T(0)};
  if (i < n && sub == 0) {
    m = mask[i];
    if (x1) di = load_vec3(dinv_c + 3 * (size_t)i);
  }
  if (i < n) {
    const int j1 = ptr[i + 1];
    // two entries per lane at a time (16 per row), loads issued together;
    // the second index is clamped to the row's last entry and skipped
    for (int j = ptr[i] + sub; j < j1; j += 16) {
      const int ja = j, jb = min(j + 8, j1 - 1);
      const int ca = col[ja], cb = col[jb];
      const int ea = wexp ? (int)wexp[ca] : 0, eb = wexp ? (int)wexp[cb] : 0;
      const Vec3<T> ua = load_vec3(rf + 3 * (size_t)ca), ub = load_vec3(rf + 3 * (size_t)cb);
      const T wa = ea ? (T)ldexpf(1.0f, 1 - ea) : (T)w[ja];
      a0 += wa * ua.a;
      a1 += wa * ua.b;
      a2 += wa * ua.c;
      if (j + 8 < j1) {
        const T wb = eb ? (T)ldexpf(1.0f, 1 - eb) : (T)w[jb];
        a0 += wb * ub.a;
        a1 += wb * ub.b;
        a2 += wb * ub.c;
      }
    }
  }
  a0 += __shfl_xor(a0, 1, 64);
  a1 += __shfl_xor(a1, 1, 64);
  a2 += __shfl_xor(a2, 1, 64);
  a0 += __shfl_xor(a0, 2, 64);
  a1 += __shfl_xor(a1, 2, 64);
  a2 += __shfl_xor(a2, 2, 64);
  a0 += __shfl_xor(a0, 4, 64);
  a1 += __shfl_xor(a1, 4, 64);
  a2 += __shfl_xor(a2, 4, 64);
  if (i < n && sub == 0) {
    const T v[3] = {(m & 1) ? T(0) : a0, (m & 2) ? T(0) : a1, (m & 4) ? T(0) : a2};
    store_vec3(rc + 3 * (size_t)i, v);
    if (x1) {  // first Chebyshev step of the coarse level from zero
      const T u[3] = {(T)c0 * di.a * v[0], (T)c0 * di.b * v[1], (T)c0 * di.c * v[2]};
      store_vec3(x1 + 3 * (size_t)i, u);
    }
  }
}

template <class T>
void launch_prolong_add(const TransferCSR &P, const uint8_t *mask_f, const T *xc, T *xf,
                        hipStream_t st, int nrhs, int64_t vsc, int64_t vsf) {
  if (P.n_rows == 0) return;
  hipLaunchKernelGGL(k_prolong_add<T>, dim3((P.n_rows + 255) / 256, nrhs), dim3(256), 0, st,
                     P.n_rows, P.ptr, P.col, P.w, mask_f, xc, xf, vsc, vsf, P.run, P.wexp);
  TMRX_LAUNCH_CHECK();
}

template <class T>
void launch_restrict(const TransferCSR &R, const uint8_t *mask_c, const T *rf, T *rc,
                     hipStream_t st, int nrhs, int64_t vsf, int64_t vsc, const T *dinv_c,
                     double c0, T *x1) {
  if (R.n_rows == 0) return;
  const int64_t threads = 8 * (int64_t)R.n_rows;
  hipLaunchKernelGGL(k_restrict<T>, dim3((unsigned)((threads + 255) / 256), nrhs), dim3(256), 0,
                     st, R.n_rows, R.ptr, R.col, R.w, mask_c, rf, rc, vsf, vsc, R.run, dinv_c, c0,
                     x1, R.wexp);
  TMRX_LAUNCH_CHECK();
}

// ---------------------------------------------------------------------------
// Chebyshev first step from a zero guess: d = c Dinv b, x = d
template <class T>
__global__ void k_cheb_first(int n, const T *__restrict__ b, const T *__restrict__ dinv, double c,
                             T *__restrict__ d, T *__restrict__ x, int64_t vs,
                             const int *__restrict__ run) {
  if (run && *run == 0) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t off = (int64_t)blockIdx.y * vs;
  const T v = (T)c * dinv[i] * b[off + i];
  if (d) d[off + i] = v;
  x[off + i] = v;
}

template <class T>
void launch_cheb_first(int ndof, const T *b, const T *dinv, double c, T *d, T *x,
                       hipStream_t st, int nrhs, const int *run) {
  hipLaunchKernelGGL(k_cheb_first<T>, dim3((ndof + 255) / 256, nrhs), dim3(256), 0, st, ndof, b,
                     dinv, c, d, x, (int64_t)ndof, run);
  TMRX_LAUNCH_CHECK();
}

// coarsest level: x = Ainv b with Ainv block diagonal (one dense FP64
// symmetric block per independent component; one wave per row).  Row i of
// the coarse vector belongs to the block whose columns start at col0[i]
// (nb[i] of them) and whose row i is stored at Ainv + aoff[i].
template <class T, class TA>
__global__ __launch_bounds__(256) void k_dense_solve(int n, const TA *__restrict__ Ainv,
                                                     const int *__restrict__ col0,
                                                     const int *__restrict__ nb,
                                                     const int64_t *__restrict__ aoff,
                                                     const T *__restrict__ b, T *__restrict__ x,
                                                     const int *__restrict__ run) {
  if (run && *run == 0) return;
  const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (row >= n) return;
  b += (int64_t)blockIdx.y * n;
  x += (int64_t)blockIdx.y * n;
  double s = 0.0;
  const TA *a = Ainv + aoff[row];
  const T *bb = b + col0[row];
  const int m = nb[row];
  // eight entries per lane in flight, added in the order of a one-at-a-time
  // loop (indices clamped to the row, extra terms skipped)
  for (int j0 = lane; j0 < m; j0 += 8 * 64) {
    double av[8], bv[8];
#pragma unroll
    for (int k = 0; k < 8; k++) {
      const int j = min(j0 + 64 * k, m - 1);
      av[k] = (double)a[j];
      bv[k] = (double)bb[j];
    }
#pragma unroll
    for (int k = 0; k < 8; k++)
      if (j0 + 64 * k < m) s += av[k] * bv[k];
  }
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  if (lane == 0) x[row] = (T)s;
}

template <class T>
void launch_dense_solve(const CoarseBlocks &cb, const T *b, T *x, hipStream_t st, int nrhs) {
  if (cb.n == 0) return;
  if (sizeof(T) == 4 && cb.Ainv32)
    hipLaunchKernelGGL((k_dense_solve<T, float>), dim3((cb.n + 3) / 4, nrhs), dim3(256), 0, st,
                       cb.n, cb.Ainv32, cb.col0, cb.nb, cb.aoff, b, x, cb.run);
  else
    hipLaunchKernelGGL((k_dense_solve<T, double>), dim3((cb.n + 3) / 4, nrhs), dim3(256), 0, st,
                       cb.n, cb.Ainv, cb.col0, cb.nb, cb.aoff, b, x, cb.run);
  TMRX_LAUNCH_CHECK();
}

template <class TI, class TO>
__global__ void k_convert(int64_t n, const TI *__restrict__ a, TO *__restrict__ b) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (TO)a[i];
}
template <class TI, class TO>
void launch_convert(int64_t n, const TI *a, TO *b, hipStream_t st) {
  if (n == 0) return;
  hipLaunchKernelGGL((k_convert<TI, TO>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n,
                     a, b);
  TMRX_LAUNCH_CHECK();
}

template <class T>
__global__ void k_fill(int64_t n, T v, T *__restrict__ a) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}
template <class T>
void launch_fill(int64_t n, T v, T *a, hipStream_t st) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_fill<T>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, v, a);
  TMRX_LAUNCH_CHECK();
}

// y[r][v] = sum_j w_j x[col_j][v] over the CSR row r (nv values per node):
// the interpolation of a nodal field from one forest to another
__global__ void k_interp_nodal(int nrows, const int *__restrict__ ptr, const int *__restrict__ col,
                               const double *__restrict__ w, int nv, const double *__restrict__ x,
                               double *__restrict__ y) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)nrows * nv) return;
  const int r = (int)(t / nv), v = (int)(t % nv);
  double s = 0.0;
  for (int j = ptr[r]; j < ptr[r + 1]; j++) s += w[j] * x[(int64_t)col[j] * nv + v];
  y[t] = s;
}

void launch_interp_nodal(int nrows, const int *ptr, const int *col, const double *w, int nv,
                         const double *x, double *y, hipStream_t st) {
  const int64_t n = (int64_t)nrows * nv;
  if (n == 0) return;
  hipLaunchKernelGGL(k_interp_nodal, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, nrows,
                     ptr, col, w, nv, x, y);
  TMRX_LAUNCH_CHECK();
}

// power-iteration start: x_i = 0.5 + U_i on free dofs (U_i in [0, 1) from
// the splitmix64 hash of seed + i, so every dof draws independently), 0 on
// clamped ones
__global__ void k_start_vector(int64_t n, const uint8_t *__restrict__ bcmask, uint64_t seed,
                               double *__restrict__ x) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint64_t z = seed + 0x9e3779b97f4a7c15ULL * (uint64_t)(i + 1);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  z ^= z >> 31;
  x[i] = ((bcmask[i / 3] >> (i % 3)) & 1) ? 0.0 : 0.5 + (double)(z >> 11) * 0x1.0p-53;
}

void launch_start_vector(int64_t n, const uint8_t *bcmask, uint64_t seed, double *x,
                         hipStream_t st) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_start_vector, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n,
                     bcmask, seed, x);
  TMRX_LAUNCH_CHECK();
}

// *bad |= 1 when an entry is not finite or lies outside [lo, hi] (a NaN
// fails both comparisons)
__global__ void k_check_range(int64_t n, const double *__restrict__ v, double lo, double hi,
                              int *bad) {
  bool b = false;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double x = v[i];
    if (!(x >= lo && x <= hi)) b = true;
  }
  if (b) atomicOr(bad, 1);
}

void launch_check_range(int64_t n, const double *v, double lo, double hi, int *bad,
                        hipStream_t st) {
  if (n == 0) return;
  const int64_t nb = std::min<int64_t>((n + 255) / 256, 2048);
  hipLaunchKernelGGL(k_check_range, dim3((unsigned)nb), dim3(256), 0, st, n, v, lo, hi, bad);
  TMRX_LAUNCH_CHECK();
}

__global__ void k_zero_clamped(int n_nodes, const uint8_t *__restrict__ bcmask,
                               double *__restrict__ v, int64_t vs) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_nodes) return;
  v += (int64_t)blockIdx.y * vs;
  const uint8_t m = bcmask[i];
  if (m & 1) v[3 * (int64_t)i] = 0.0;
  if (m & 2) v[3 * (int64_t)i + 1] = 0.0;
  if (m & 4) v[3 * (int64_t)i + 2] = 0.0;
}

__global__ void k_axpy_bcast(int64_t n, double alpha, const double *__restrict__ y,
                             double *__restrict__ v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) v[(int64_t)blockIdx.y * n + i] += alpha * y[i];
}

void launch_axpy_bcast(int64_t n, int nrhs, double alpha, const double *y, double *v,
                       hipStream_t st) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_axpy_bcast, dim3((unsigned)((n + 255) / 256), nrhs), dim3(256), 0, st, n,
                     alpha, y, v);
  TMRX_LAUNCH_CHECK();
}

void launch_zero_clamped(int n_nodes, const uint8_t *bcmask, double *v, hipStream_t st, int nrhs,
                         int64_t vs) {
  if (n_nodes == 0) return;
  hipLaunchKernelGGL(k_zero_clamped, dim3((n_nodes + 255) / 256, nrhs), dim3(256), 0, st, n_nodes,
                     bcmask, v, vs);
  TMRX_LAUNCH_CHECK();
}

// matrix-core records carry their element's FP32 stiffness scale in word 4
__global__ void k_rec_scale(int nrec, const int *__restrict__ rec_elem,
                            const float *__restrict__ s32, uint32_t *__restrict__ rec_words) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nrec) return;
  rec_words[6 * (size_t)i + 4] = __float_as_uint(s32[rec_elem[i]]);
}

void launch_rec_scale(int nrec, const int *rec_elem, const float *s32, uint32_t *rec_words,
                      hipStream_t st) {
  if (nrec == 0) return;
  hipLaunchKernelGGL(k_rec_scale, dim3((nrec + 255) / 256), dim3(256), 0, st, nrec, rec_elem, s32,
                     rec_words);
  TMRX_LAUNCH_CHECK();
}

// ---------------------------------------------------------------------------
// SIMP / RAMP penalty (src/topology/TMROctConstitutive.cpp:498-551) times
// the element size factor 2^-level.
__global__ void k_simp_scale(int n, const double *__restrict__ rho, const int16_t *__restrict__ lev,
                             double q, double k0, int ptype, double *__restrict__ s64,
                             float *__restrict__ s32) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double r = rho[i];
  const double pen = (ptype == 0) ? pow(r, q) + k0 : r / (1.0 + q * (1.0 - r)) + k0;
  const double s = ldexp(pen, -(int)lev[i]);
  s64[i] = s;
  s32[i] = (float)s;
}

void launch_simp_scale(int n, const double *rho, const int16_t *level, double q, double k0,
                       int ptype, double *s64, float *s32, hipStream_t st) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_simp_scale, dim3((n + 255) / 256), dim3(256), 0, st, n, rho, level, q, k0,
                     ptype, s64, s32);
  TMRX_LAUNCH_CHECK();
}

__global__ void k_restrict_scale(int nc, const int *__restrict__ ptr, const int *__restrict__ child,
                                 const float *__restrict__ vw, const double *__restrict__ sf,
                                 const int16_t *__restrict__ lf, const int16_t *__restrict__ lc,
                                 double *__restrict__ s64, float *__restrict__ s32) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nc) return;
  double m = 0.0;
  for (int j = ptr[i]; j < ptr[i + 1]; j++) {
    const int c = child[j];
    m += (double)vw[j] * ldexp(sf[c], (int)lf[c]);  // material factor of the child
  }
  const double s = ldexp(m, -(int)lc[i]);
  s64[i] = s;
  s32[i] = (float)s;
}

void launch_restrict_scale(int nc, const int *ptr, const int *child, const float *vw,
                           const double *s_fine, const int16_t *lev_fine, const int16_t *lev_coarse,
                           double *s64, float *s32, hipStream_t st) {
  if (nc == 0) return;
  hipLaunchKernelGGL(k_restrict_scale, dim3((nc + 255) / 256), dim3(256), 0, st, nc, ptr, child, vw,
                     s_fine, lev_fine, lev_coarse, s64, s32);
  TMRX_LAUNCH_CHECK();
}

// ---------------------------------------------------------------------------
// nodal design variables (one per node, TMROctConstitutive with nvars = 1)

// trilinear shape function of corner a (bits x, y, z) at Gauss point g (bit
// d of g set: coordinate d at +1/sqrt(3)): each factor is (1 +- 1/sqrt(3))/2
__device__ __forceinline__ double gauss_N(int a, int g) {
  const double gp = 0.57735026918962576451;  // 1/sqrt(3)
  const double hi = 0.5 * (1.0 + gp), lo = 0.5 * (1.0 - gp);
  const int same = ~(a ^ g);
  return ((same & 1) ? hi : lo) * ((same & 2) ? hi : lo) * ((same & 4) ? hi : lo);
}

__device__ __forceinline__ double node_value(int g, const double *__restrict__ xn,
                                             const int *__restrict__ dep_ptr,
                                             const int *__restrict__ dep_conn,
                                             const float *__restrict__ dep_w) {
  if (g >= 0) return xn[g];
  const int d = -g - 1;
  double v = 0.0;
  for (int j = dep_ptr[d]; j < dep_ptr[d + 1]; j++) v += (double)dep_w[j] * xn[dep_conn[j]];
  return v;
}

__device__ __forceinline__ double penalty_of(double r, double q, double k0, int ptype) {
  return (ptype == 0) ? pow(r, q) + k0 : r / (1.0 + q * (1.0 - r)) + k0;
}

__device__ __forceinline__ double dpenalty_of(double r, double q, int ptype) {
  if (ptype == 0) return (q > 1.0) ? q * pow(r, q - 1.0) : 1.0;
  const double t = 1.0 + q * (1.0 - r);
  return (q + 1.0) / (t * t);
}

__global__ void k_nodal_scale(int ne, const int *__restrict__ conn, const int *__restrict__ dep_ptr,
                              const int *__restrict__ dep_conn, const float *__restrict__ dep_w,
                              const double *__restrict__ xn, const int16_t *__restrict__ lev,
                              double q, double k0, int ptype, double *__restrict__ rho_gp,
                              double *__restrict__ s64, float *__restrict__ s32) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ne) return;
  double xa[8];
#pragma unroll
  for (int a = 0; a < 8; a++) xa[a] = node_value(conn[8 * (size_t)e + a], xn, dep_ptr, dep_conn, dep_w);
  const int l = lev[e];
#pragma unroll
  for (int g = 0; g < 8; g++) {
    double r = 0.0;
#pragma unroll
    for (int a = 0; a < 8; a++) r += gauss_N(a, g) * xa[a];
    const double sc = ldexp(penalty_of(r, q, k0, ptype), -l);
    const size_t k = 8 * (size_t)e + g;
    if (rho_gp) rho_gp[k] = r;
    s64[k] = sc;
    s32[k] = (float)sc;
  }
}

void launch_nodal_scale(int nelem, const int *conn, const int *dep_ptr, const int *dep_conn,
                        const float *dep_w, const double *xn, const int16_t *level, double q,
                        double k0, int penalty_type, double *rho_gp, double *s64, float *s32,
                        hipStream_t st) {
  if (nelem == 0) return;
  hipLaunchKernelGGL(k_nodal_scale, dim3((nelem + 255) / 256), dim3(256), 0, st, nelem, conn,
                     dep_ptr, dep_conn, dep_w, xn, level, q, k0, penalty_type, rho_gp, s64, s32);
  TMRX_LAUNCH_CHECK();
}

__global__ void k_weight_restrict(int nc, const int *__restrict__ ptr, const int *__restrict__ col,
                                  const float *__restrict__ w, const double *__restrict__ xf,
                                  double *__restrict__ xc) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nc) return;
  double s = 0.0, ws = 0.0;
  for (int j = ptr[i]; j < ptr[i + 1]; j++) {
    const double wj = (double)w[j];
    s += wj * xf[col[j]];
    ws += wj;
  }
  xc[i] = s / ws;
}

void launch_weight_restrict(const TransferCSR &R, const double *xf, double *xc, hipStream_t st) {
  if (R.n_rows == 0) return;
  hipLaunchKernelGGL(k_weight_restrict, dim3((R.n_rows + 255) / 256), dim3(256), 0, st, R.n_rows,
                     R.ptr, R.col, R.w, xf, xc);
  TMRX_LAUNCH_CHECK();
}

// grad u at Gauss point g of an element (H[3 c + d] = du_c / dx_d)
__device__ __forceinline__ void grad_at(const double *__restrict__ G, const double ue[24],
                                        double H[9]) {
  for (int i = 0; i < 9; i++) H[i] = 0.0;
  for (int a = 0; a < 8; a++)
    for (int c = 0; c < 3; c++)
      for (int d = 0; d < 3; d++) H[3 * c + d] += ue[3 * a + c] * G[3 * a + d];
}

__global__ void k_nodal_energy(int ne, const int *__restrict__ conn,
                               const int *__restrict__ dep_ptr, const int *__restrict__ dep_conn,
                               const float *__restrict__ dep_w, const uint8_t *__restrict__ bcmask,
                               const int *__restrict__ cls, const double *__restrict__ Gtab,
                               const double *__restrict__ u, const double *__restrict__ v,
                               double *__restrict__ out) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ne) return;
  double ue[24], ve[24];
  for (int a = 0; a < 8; a++) {
    double t[3];
    load_node<double>(conn[8 * (size_t)e + a], u, dep_ptr, dep_conn, dep_w, bcmask, t);
    ue[3 * a] = t[0];
    ue[3 * a + 1] = t[1];
    ue[3 * a + 2] = t[2];
    load_node<double>(conn[8 * (size_t)e + a], v, dep_ptr, dep_conn, dep_w, bcmask, t);
    ve[3 * a] = t[0];
    ve[3 * a + 1] = t[1];
    ve[3 * a + 2] = t[2];
  }
  const double *Gt = Gtab + NODAL_TABLE * (size_t)cls[e];
  const double lam = Gt[192], mu = Gt[193];
  for (int g = 0; g < 8; g++) {
    double Hu[9], Hv[9];
    grad_at(Gt + 24 * g, ue, Hu);
    grad_at(Gt + 24 * g, ve, Hv);
    // eps_v : C eps_u with symmetric strains eps = (H + H^T) / 2
    double se = 0.0;
    for (int c = 0; c < 3; c++)
      for (int d = 0; d < 3; d++)
        se += 0.25 * (Hu[3 * c + d] + Hu[3 * d + c]) * (Hv[3 * c + d] + Hv[3 * d + c]);
    const double tru = Hu[0] + Hu[4] + Hu[8], trv = Hv[0] + Hv[4] + Hv[8];
    out[8 * (size_t)e + g] = lam * tru * trv + 2.0 * mu * se;
  }
}

void launch_nodal_energy(int nelem, const int *conn, const int *dep_ptr, const int *dep_conn,
                         const float *dep_w, const uint8_t *bcmask, const int *cls,
                         const double *G, const double *u, const double *v, double *out,
                         hipStream_t st) {
  if (nelem == 0) return;
  hipLaunchKernelGGL(k_nodal_energy, dim3((nelem + 255) / 256), dim3(256), 0, st, nelem, conn,
                     dep_ptr, dep_conn, dep_w, bcmask, cls, G, u, v, out);
  TMRX_LAUNCH_CHECK();
}

__global__ __launch_bounds__(WG_ELEMS) void k_nodal_grad(
    const int *__restrict__ wg_node_ptr, const int *__restrict__ wg_elem_ptr,
    const uint4 *__restrict__ lconn, const uint8_t *__restrict__ color, int ncolors,
    const int16_t *__restrict__ lev, const double *__restrict__ rho_gp, double q, int ptype,
    const double *__restrict__ e_gp, const double *__restrict__ ex_gp,
    double *__restrict__ partial) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int wg = blockIdx.x, tid = threadIdx.x;
  const int nb = wg_node_ptr[wg], nn = wg_node_ptr[wg + 1] - nb;
  double *acc = reinterpret_cast<double *>(smem_raw);
  for (int i = tid; i < nn; i += WG_ELEMS) acc[i] = 0.0;
  __syncthreads();
  const int e0 = wg_elem_ptr[wg], ne = wg_elem_ptr[wg + 1] - e0;
  int lc[8];
  int col = -1;
  double t[8];
  if (tid < ne) {
    const int e = e0 + tid;
    const uint4 qq = lconn[e];
    lc[0] = qq.x & 0xffff; lc[1] = qq.x >> 16;
    lc[2] = qq.y & 0xffff; lc[3] = qq.y >> 16;
    lc[4] = qq.z & 0xffff; lc[5] = qq.z >> 16;
    lc[6] = qq.w & 0xffff; lc[7] = qq.w >> 16;
    col = color[e];
    double wgp[8];
    for (int g = 0; g < 8; g++) {
      const size_t k = 8 * (size_t)e + g;
      const double dp = ldexp(dpenalty_of(rho_gp[k], q, ptype), -(int)lev[e]);
      wgp[g] = (ex_gp ? ex_gp[k] : 0.0) - dp * e_gp[k];
    }
    for (int a = 0; a < 8; a++) {
      double v = 0.0;
      for (int g = 0; g < 8; g++) v += gauss_N(a, g) * wgp[g];
      t[a] = v;
    }
  }
  for (int c = 0; c < ncolors; c++) {
    if (col == c) {
      for (int a = 0; a < 8; a++) acc[lc[a]] += t[a];
    }
    __syncthreads();
  }
  for (int i = tid; i < nn; i += WG_ELEMS) {
    partial[3 * (size_t)(nb + i)] = acc[i];
    partial[3 * (size_t)(nb + i) + 1] = 0.0;
    partial[3 * (size_t)(nb + i) + 2] = 0.0;
  }
}

void launch_nodal_grad(const ElemBatches &eb, const int16_t *lev, const double *rho_gp, double q,
                       int penalty_type, const double *e_gp, const double *ex_gp,
                       double *partial, hipStream_t st) {
  if (eb.n_wg == 0) return;
  const size_t sh = (size_t)eb.max_wg_nodes * sizeof(double);
  hipLaunchKernelGGL(k_nodal_grad, dim3(eb.n_wg), dim3(WG_ELEMS), sh, st, eb.wg_node_ptr,
                     eb.wg_elem_ptr, reinterpret_cast<const uint4 *>(eb.lconn), eb.color,
                     eb.ncolors, lev, rho_gp, q, penalty_type, e_gp, ex_gp, partial);
  TMRX_LAUNCH_CHECK();
}

__global__ void k_take_comp0(int n, const double *__restrict__ in, double *__restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = in[3 * (size_t)i];
}

void launch_take_comp0(int n, const double *in, double *out, hipStream_t st) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_take_comp0, dim3((n + 255) / 256), dim3(256), 0, st, n, in, out);
  TMRX_LAUNCH_CHECK();
}

// ---------------------------------------------------------------------------
// FP64 PCG vector kernels with block partial reductions (fixed grid, so the
// reduction order, and the result, is the same on every run)
constexpr int RED_BLOCK = 256;
constexpr int RED_MAX_BLOCKS = 2048;

int reduce_blocks(int64_t n) {
  int64_t b = (n + RED_BLOCK - 1) / RED_BLOCK;
  return (int)(b < RED_MAX_BLOCKS ? (b < 1 ? 1 : b) : RED_MAX_BLOCKS);
}

__global__ __launch_bounds__(RED_BLOCK) void k_pcg_update_xr(int64_t n, double *__restrict__ x,
                                                             double *__restrict__ r,
                                                             const double *__restrict__ p,
                                                             const double *__restrict__ q,
                                                             const double *__restrict__ sc,
                                                             double *__restrict__ partials,
                                                             float *__restrict__ r32) {
  __shared__ double red[RED_BLOCK / 64];
  const double alpha = sc[SC_ALPHA];
  double acc = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * RED_BLOCK + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * RED_BLOCK) {
    x[i] += alpha * p[i];
    const double rn = r[i] - alpha * q[i];
    r[i] = rn;
    if (r32) r32[i] = (float)rn;
    acc += rn * rn;
  }
  const double s = block_sum(acc, red);
  if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

void launch_pcg_update_xr(int64_t ndof, double *x, double *r, const double *p, const double *q,
                          const double *sc, double *partials, float *r32, hipStream_t st) {
  hipLaunchKernelGGL(k_pcg_update_xr, dim3(reduce_blocks(ndof)), dim3(RED_BLOCK), 0, st, ndof, x, r,
                     p, q, sc, partials, r32);
  TMRX_LAUNCH_CHECK();
}

__global__ __launch_bounds__(RED_BLOCK) void k_dot2(int64_t n, const double *__restrict__ z,
                                                    const double *__restrict__ r,
                                                    const double *__restrict__ q,
                                                    double *__restrict__ partials) {
  __shared__ double red[RED_BLOCK / 64];
  double a = 0.0, b = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * RED_BLOCK + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * RED_BLOCK) {
    a += z[i] * r[i];
    b += z[i] * q[i];
  }
  const double sa = block_sum(a, red);
  __syncthreads();
  const double sb = block_sum(b, red);
  if (threadIdx.x == 0) {
    partials[2 * blockIdx.x] = sa;
    partials[2 * blockIdx.x + 1] = sb;
  }
}

void launch_dot2(int64_t ndof, const double *z, const double *r, const double *q,
                 double *partials, hipStream_t st) {
  hipLaunchKernelGGL(k_dot2, dim3(reduce_blocks(ndof)), dim3(RED_BLOCK), 0, st, ndof, z, r, q,
                     partials);
  TMRX_LAUNCH_CHECK();
}

__global__ __launch_bounds__(RED_BLOCK) void k_dot1(int64_t n, const double *__restrict__ a,
                                                    const double *__restrict__ b,
                                                    double *__restrict__ partials) {
  __shared__ double red[RED_BLOCK / 64];
  double s = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * RED_BLOCK + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * RED_BLOCK)
    s += a[i] * b[i];
  const double t = block_sum(s, red);
  if (threadIdx.x == 0) partials[blockIdx.x] = t;
}

void launch_dot1(int64_t ndof, const double *a, const double *b, double *partials,
                 hipStream_t st) {
  hipLaunchKernelGGL(k_dot1, dim3(reduce_blocks(ndof)), dim3(RED_BLOCK), 0, st, ndof, a, b,
                     partials);
  TMRX_LAUNCH_CHECK();
}

__global__ void k_pcg_update_p(int64_t n, double *__restrict__ p, const double *__restrict__ z,
                               const double *__restrict__ sc) {
  const double beta = sc[SC_BETA];
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = z[i] + beta * p[i];
}

void launch_pcg_update_p(int64_t ndof, double *p, const double *z, const double *sc,
                         hipStream_t st) {
  hipLaunchKernelGGL(k_pcg_update_p, dim3(reduce_blocks(ndof)), dim3(RED_BLOCK), 0, st, ndof, p, z,
                     sc);
  TMRX_LAUNCH_CHECK();
}

// one block sums the partials (fixed order) and updates the PCG scalars
__global__ __launch_bounds__(256) void k_finalize(int np, const double *__restrict__ part,
                                                  int mode, double *__restrict__ sc) {
  __shared__ double red[4];
  const int stride = (mode == 3 || mode == 5) ? 2 : 1;
  double a = 0.0, b = 0.0;
  for (int i = threadIdx.x; i < np; i += 256) {
    a += part[stride * i];
    if (mode == 3 || mode == 5) b += part[2 * i + 1];
  }
  const double sa = block_sum(a, red);
  __syncthreads();
  const double sb = block_sum(b, red);
  if (threadIdx.x != 0) return;
  switch (mode) {
    case 0: sc[SC_RZ] = sa; break;
    case 1:
      sc[SC_PQ] = sa;
      sc[SC_ALPHA] = sc[SC_RZ] / sa;
      break;
    case 2: sc[SC_RR] = sa; break;
    case 3:  // flexible (Polak-Ribiere) beta = z_{k+1}.(r_{k+1}-r_k) / z_k.r_k
      sc[SC_RZNEW] = sa;
      sc[SC_ZQ] = sb;
      sc[SC_BETA] = -sc[SC_ALPHA] * sb / sc[SC_RZ];
      sc[SC_RZ] = sa;
      break;
    case 4: sc[SC_BB] = sa; break;
    case 5:  // power iteration: |Dinv A x|^2 and |x|^2
      sc[SC_ZQ] = sa;
      sc[SC_RZNEW] = sb;
      break;
  }
}

void launch_finalize(int nparts, const double *partials, int mode, double *sc, hipStream_t st) {
  hipLaunchKernelGGL(k_finalize, dim3(1), dim3(256), 0, st, nparts, partials, mode, sc);
  TMRX_LAUNCH_CHECK();
}

// ---------------------------------------------------------------------------
// element strain energy u_e^T K0 u_e (unscaled) for the adjoint / compliance
// sensitivities (src/topology/TMROctConstitutive.cpp:557-654 restricted to
// element-constant density).
__global__ void k_elem_energy(int ne, const int *__restrict__ conn, const int *__restrict__ dep_ptr,
                              const int *__restrict__ dep_conn, const float *__restrict__ dep_w,
                              const uint8_t *__restrict__ bcmask, const int *__restrict__ cls,
                              const double *__restrict__ K0, const double *__restrict__ u,
                              double *__restrict__ energy) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ne) return;
  double ue[24];
  for (int a = 0; a < 8; a++) {
    double v[3];
    load_node<double>(conn[8 * e + a], u, dep_ptr, dep_conn, dep_w, bcmask, v);
    ue[3 * a] = v[0];
    ue[3 * a + 1] = v[1];
    ue[3 * a + 2] = v[2];
  }
  const double *K = K0 + 576 * cls[e];
  double s = 0.0;
  for (int i = 0; i < 24; i++) {
    double t = 0.0;
    for (int j = 0; j < 24; j++) t += K[24 * i + j] * ue[j];
    s += ue[i] * t;
  }
  energy[e] = s;
}

void launch_elem_energy(int nelem, const int *conn, const int *dep_ptr, const int *dep_conn,
                        const float *dep_w, const uint8_t *bcmask, const int *cls,
                        const double *K0, const double *u, double *energy, hipStream_t st) {
  if (nelem == 0) return;
  hipLaunchKernelGGL(k_elem_energy, dim3((nelem + 255) / 256), dim3(256), 0, st, nelem, conn,
                     dep_ptr, dep_conn, dep_w, bcmask, cls, K0, u, energy);
  TMRX_LAUNCH_CHECK();
}

// d compliance / d rho_e = -dpenalty(rho_e) 2^-level u_e^T K0 u_e
__global__ void k_compliance_grad(int n, const double *__restrict__ rho, const int16_t *__restrict__ lev,
                                  double q, int ptype, const double *__restrict__ energy,
                                  double *__restrict__ g) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double r = rho[i];
  double dp;
  if (ptype == 0) {
    dp = (q > 1.0) ? q * pow(r, q - 1.0) : 1.0;
  } else {
    const double t = 1.0 + q * (1.0 - r);
    dp = (q + 1.0) / (t * t);
  }
  g[i] = -ldexp(dp, -(int)lev[i]) * energy[i];
}

void launch_compliance_grad(int n, const double *rho, const int16_t *lev, double q, int ptype,
                            const double *energy, double *g, hipStream_t st) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_compliance_grad, dim3((n + 255) / 256), dim3(256), 0, st, n, rho, lev, q,
                     ptype, energy, g);
  TMRX_LAUNCH_CHECK();
}

// power iteration step for lambda_max(Dinv A): y <- Dinv y, partials of
// y.y and x.x
__global__ __launch_bounds__(RED_BLOCK) void k_dinv_dots(int64_t n, const double *__restrict__ dinv,
                                                         double *__restrict__ y,
                                                         const double *__restrict__ x,
                                                         double *__restrict__ partials) {
  __shared__ double red[RED_BLOCK / 64];
  double a = 0.0, b = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * RED_BLOCK + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * RED_BLOCK) {
    const double v = dinv[i] * y[i];
    y[i] = v;
    a += v * v;
    b += x[i] * x[i];
  }
  const double sa = block_sum(a, red);
  __syncthreads();
  const double sb = block_sum(b, red);
  if (threadIdx.x == 0) {
    partials[2 * blockIdx.x] = sa;
    partials[2 * blockIdx.x + 1] = sb;
  }
}

void launch_dinv_dots(int64_t n, const double *dinv, double *y, const double *x,
                      double *partials, hipStream_t st) {
  hipLaunchKernelGGL(k_dinv_dots, dim3(reduce_blocks(n)), dim3(RED_BLOCK), 0, st, n, dinv, y, x,
                     partials);
  TMRX_LAUNCH_CHECK();
}

// ---------------------------------------------------------------------------
// owner-tile apply: one pass, final values of the owned nodes, fused epilogue

// stage x of the tile's slots into LDS (owned nodes in place, halo nodes with
// hanging-node interpolation, clamped dofs zero) and zero the accumulators;
// S floats per node (S = 4 pads each node to 16 bytes with a zero)
template <class T, int S, int NT = TILE_THREADS, int SY = S>
__device__ __forceinline__ void tile_stage(int n0, int nown, const int *__restrict__ ext_l, int next,
                                           int nacc, const int *__restrict__ dep_ptr,
                                           const int *__restrict__ dep_conn,
                                           const float *__restrict__ dep_w,
                                           const uint8_t *__restrict__ bcmask,
                                           const T *__restrict__ x, T *xs, T *ys) {
  const int tid = threadIdx.x;
  constexpr int VN = 2, VH = 2;
  auto put = [&](int pos, T a, T b, T c) {
    const int q = S * tile_pos(pos);
    xs[q] = a;
    xs[q + 1] = b;
    xs[q + 2] = c;
  };
  // halo node g of slot nown + i: independent nodes load directly (masked),
  // hanging ones are interpolated from their parents
  auto put_halo = [&](int i, int g, const Vec3<T> &w, uint8_t m) {
    T v[3] = {(m & 1) ? T(0) : w.a, (m & 2) ? T(0) : w.b, (m & 4) ? T(0) : w.c};
    if (g < 0) load_node<T>(g, x, dep_ptr, dep_conn, dep_w, bcmask, v);
    put(nown + i, v[0], v[1], v[2]);
  };
  // First rounds: the halo node indices, then the owned values (one 12-byte
  // load and one mask byte per node) and the halo values all go out before
  // any is used; indices are clamped into the tile, so the loads are
  // unconditional.
  int gf[VH];
#pragma unroll
  for (int k = 0; k < VH; k++) gf[k] = next > 0 ? ext_l[min(tid + k * NT, next - 1)] : 0;
  Vec3<T> vo[VN], vh[VH];
  uint8_t mo[VN], mh[VH];
#pragma unroll
  for (int k = 0; k < VN; k++) {
    const int j = min(tid + k * NT, nown - 1);
    vo[k] = load_vec3(x + 3 * (size_t)(n0 + j));
    mo[k] = bcmask[n0 + j];
  }
#pragma unroll
  for (int k = 0; k < VH; k++) {
    const int gg = gf[k] >= 0 ? gf[k] : 0;
    vh[k] = load_vec3(x + 3 * (size_t)gg);
    mh[k] = bcmask[gg];
  }
#pragma unroll
  for (int k = 0; k < VN; k++)
    if (tid + k * NT < nown)
      put(tid + k * NT, (mo[k] & 1) ? T(0) : vo[k].a, (mo[k] & 2) ? T(0) : vo[k].b,
          (mo[k] & 4) ? T(0) : vo[k].c);
#pragma unroll
  for (int k = 0; k < VH; k++)
    if (tid + k * NT < next) put_halo(tid + k * NT, gf[k], vh[k], mh[k]);
  // further rounds (tiles of more than two nodes or halo slots per thread)
  for (int j0 = tid + VN * NT; j0 < nown; j0 += VN * NT) {
#pragma unroll
    for (int k = 0; k < VN; k++) {
      const int j = min(j0 + k * NT, nown - 1);
      vo[k] = load_vec3(x + 3 * (size_t)(n0 + j));
      mo[k] = bcmask[n0 + j];
    }
#pragma unroll
    for (int k = 0; k < VN; k++)
      if (j0 + k * NT < nown)
        put(j0 + k * NT, (mo[k] & 1) ? T(0) : vo[k].a, (mo[k] & 2) ? T(0) : vo[k].b,
            (mo[k] & 4) ? T(0) : vo[k].c);
  }
  for (int i0 = tid + VH * NT; i0 < next; i0 += VH * NT) {
    int g[VH];
#pragma unroll
    for (int k = 0; k < VH; k++) g[k] = ext_l[min(i0 + k * NT, next - 1)];
#pragma unroll
    for (int k = 0; k < VH; k++) {
      const int gg = g[k] >= 0 ? g[k] : 0;
      vh[k] = load_vec3(x + 3 * (size_t)gg);
      mh[k] = bcmask[gg];
    }
#pragma unroll
    for (int k = 0; k < VH; k++)
      if (i0 + k * NT < next) put_halo(i0 + k * NT, g[k], vh[k], mh[k]);
  }
  if (S == 4)
    for (int i = tid; i < ((nown + next + 7) & ~7); i += NT) xs[4 * i + 3] = T(0);
  // accumulators: SY * round8(nacc) values, a multiple of 16 bytes starting
  // 16-byte aligned, cleared 16 bytes at a time
  uint4 *y16 = reinterpret_cast<uint4 *>(ys);
  const int n16 = (int)((SY * ((nacc + 7) & ~7) * sizeof(T)) / 16);
  for (int i = tid; i < n16; i += NT) y16[i] = make_uint4(0u, 0u, 0u, 0u);
}

// owned rows: add the hanging-node sums, then the mode's epilogue
// x of the bf16 planes (hi, mid, lo: 4 bf16 per node each, planes pstride
// nodes apart): hi + mid + lo restores the FP32 value exactly
__device__ __forceinline__ float plane_get(uint2 w, int c) {
  const uint32_t u = c == 0 ? (w.x << 16) : c == 1 ? (w.x & 0xffff0000u) : (w.y << 16);
  return __uint_as_float(u);
}

// the global inputs of one owned node's epilogue
template <class T>
struct EpiIn {
  Vec3<T> b, d, w, x;
  uint8_t m;
};
template <class T, int MODE>
__device__ __forceinline__ EpiIn<T> epi_load(int node, const uint8_t *__restrict__ bcmask,
                                             const T *__restrict__ x, const T *__restrict__ b,
                                             const T *__restrict__ d, const T *__restrict__ dinv,
                                             int cflags) {
  const size_t k3 = 3 * (size_t)node;
  EpiIn<T> e{{T(0), T(0), T(0)}, {T(0), T(0), T(0)}, {T(0), T(0), T(0)}, {T(0), T(0), T(0)}, 0};
  if (MODE == ASM_RESID || MODE == ASM_CHEB) e.b = load_vec3(b + k3);
  if (MODE == ASM_CHEB) {
    if (cflags & CHEB_D_READ) e.d = load_vec3(d + k3);
    e.w = load_vec3(dinv + k3);
  }
  if (MODE == ASM_STORE || MODE == ASM_DOT) e.x = load_vec3(x + k3);
  e.m = bcmask[node];
  return e;
}

template <class T, int MODE, int S, int NT = TILE_THREADS, bool SPLIT = false, int SX = S>
__device__ __forceinline__ void tile_finish(
    int t, int n0, int nown, const int *__restrict__ comb_ptr, const int *__restrict__ comb_slot,
    const float *__restrict__ comb_w, const uint8_t *__restrict__ bcmask, const T *xs, const T *ys,
    const T *__restrict__ x, const T *__restrict__ b, T *__restrict__ y, T *__restrict__ x_out,
    T *__restrict__ d, const T *__restrict__ dinv, double ca, double cc, int cflags,
    double *__restrict__ partials, double *red, int pstride = 0, bool hang = true) {
  auto xat = [&](int pos, int c) -> T {
    if constexpr (SPLIT) {
      const uint2 *ph = reinterpret_cast<const uint2 *>(xs);
      return (plane_get(ph[pos], c) + plane_get(ph[pos + pstride], c)) +
             plane_get(ph[pos + 2 * pstride], c);
    } else {
      return xs[SX * pos + c];
    }
  };
  double dot = 0.0;
  // two owned nodes per thread at a time: both nodes' loads (one per node
  // and vector, indices clamped into the tile) go out before either is used
  for (int i0 = threadIdx.x; i0 < nown; i0 += 2 * NT) {
    EpiIn<T> ein[2];
#pragma unroll
    for (int u = 0; u < 2; u++)
      ein[u] = epi_load<T, MODE>(n0 + min(i0 + u * NT, nown - 1), bcmask, x, b, d, dinv, cflags);
#pragma unroll
    for (int u = 0; u < 2; u++) {
      const int i = i0 + u * NT;
      if (i >= nown) break;
      const int node = n0 + i;
      const int ti = tile_pos(i);
      const int pi = S * ti;
      const size_t k3 = 3 * (size_t)node;
      const EpiIn<T> &e = ein[u];
      const T bv[3] = {e.b.a, e.b.b, e.b.c}, dv[3] = {e.d.a, e.d.b, e.d.c};
      const T wv[3] = {e.w.a, e.w.b, e.w.c}, xv[3] = {e.x.a, e.x.b, e.x.c};
      const uint8_t m = e.m;
      T a[3] = {ys[pi], ys[pi + 1], ys[pi + 2]};
      const int j0 = hang ? comb_ptr[node] : 0, j1 = hang ? comb_ptr[node + 1] : 0;
      for (int j = j0; j < j1; j++) {
        const int sl = comb_slot[j];
        const T w = (T)comb_w[j];
        a[0] += w * ys[S * sl];
        a[1] += w * ys[S * sl + 1];
        a[2] += w * ys[S * sl + 2];
      }
      T yo[3], dno[3];
#pragma unroll
      for (int c = 0; c < 3; c++) {
        const bool bc = (m >> c) & 1;
        if (MODE == ASM_STORE) {
          yo[c] = bc ? xv[c] : a[c];
        } else if (MODE == ASM_DOT) {
          const T xx = bc ? xv[c] : xat(ti, c);
          const T v = bc ? xx : a[c];
          yo[c] = v;
          dot += (double)xx * (double)v;
        } else if (MODE == ASM_RESID) {
          yo[c] = bc ? T(0) : bv[c] - a[c];
        } else if (MODE == ASM_CHEB) {
          if (bc) {
            dno[c] = T(0);
            yo[c] = T(0);
          } else {
            const T r = bv[c] - a[c];
            const T xc = xat(ti, c);
            const T dp = (cflags & CHEB_D_IS_X) ? xc : dv[c];
            const T dn = (T)ca * dp + (T)cc * wv[c] * r;
            dno[c] = dn;
            yo[c] = xc + dn;
          }
        }
      }
      // one store per node and vector
      if (MODE == ASM_CHEB) {
        if (cflags & CHEB_D_STORE) store_vec3(d + k3, dno);
        store_vec3(x_out + k3, yo);
      } else {
        store_vec3(y + k3, yo);
      }
    }
  }
  if (MODE == ASM_DOT) {
    const double sum = block_sum(dot, red);
    if (threadIdx.x == 0) partials[t] = sum;
  }
}

// 8-point Walsh-Hadamard transform over the corners (corner a = bits x, y, z):
// v[t] <- sum_a (-1)^popcount(t & a) v[a]
template <class T>
__device__ __forceinline__ void wht8(T v[8]) {
#pragma unroll
  for (int b = 1; b < 8; b <<= 1)
#pragma unroll
    for (int a = 0; a < 8; a++)
      if (!(a & b)) {
        const T u = v[a], w = v[a | b];
        v[a] = u + w;
        v[a | b] = u - w;
      }
}

// Element product y_e = s K0 x_e through the reflection symmetry of a box
// element: K0 commutes with the three coordinate reflections (corner bit d
// flipped, displacement component d negated), so after a Walsh-Hadamard
// transform of each displacement component over the corners it is
// block-diagonal -- for every character s in {0..7} one 3x3 block coupling
// X_0[s^1], X_1[s^2], X_2[s^4].  M holds the 8 blocks (72 entries, the
// transform normalisation 1/64 folded in): 72 FMAs and 144 add/subtracts
// in place of the 576 FMAs of the dense product.  With SCALED false the
// product is K0 x_e, and the caller applies s in its accumulation FMAs.
template <class T, bool SCALED = true>
__device__ __forceinline__ void refl_elem(const T *__restrict__ M, T sc, const T xe[24], T ye[24]) {
  T X0[8], X1[8], X2[8];
#pragma unroll
  for (int a = 0; a < 8; a++) {
    X0[a] = xe[3 * a];
    X1[a] = xe[3 * a + 1];
    X2[a] = xe[3 * a + 2];
  }
  wht8(X0);
  wht8(X1);
  wht8(X2);
  T Y0[8], Y1[8], Y2[8];
#pragma unroll
  for (int s = 0; s < 8; s++) {
    const T *__restrict__ m = M + 9 * s;
    const T v[3] = {SCALED ? sc * X0[s ^ 1] : X0[s ^ 1], SCALED ? sc * X1[s ^ 2] : X1[s ^ 2],
                    SCALED ? sc * X2[s ^ 4] : X2[s ^ 4]};
    T o[3];
    // only the structural non-zeros (compile-time pattern, 45 of 72)
#pragma unroll
    for (int r = 0; r < 3; r++) {
      T acc = T(0);
#pragma unroll
      for (int c = 0; c < 3; c++)
        if (refl_nz(s, 3 * r + c)) acc += m[3 * r + c] * v[c];
      o[r] = acc;
    }
    Y0[s ^ 1] = o[0];
    Y1[s ^ 2] = o[1];
    Y2[s ^ 4] = o[2];
  }
  wht8(Y0);
  wht8(Y1);
  wht8(Y2);
#pragma unroll
  for (int a = 0; a < 8; a++) {
    ye[3 * a] = Y0[a];
    ye[3 * a + 1] = Y1[a];
    ye[3 * a + 2] = Y2[a];
  }
}

// Nodal-design element product y_e = sum_g s_g K_g x_e in gradient form: at
// every Gauss point H = grad u, sigma = s_g (lambda tr(H) I + mu (H + H^T))
// (lambda, mu carry the Gauss weight |J|), y_a += sigma grad N_a; only the
// rows of the corners in `mask` are formed.
template <class T>
__device__ __forceinline__ void nodal_elem(const T *__restrict__ Gt, const T *__restrict__ sg,
                                           const T xe[24], int mask, T ye[24]) {
  const T lam = Gt[192], mu = Gt[193];
#pragma unroll
  for (int i = 0; i < 24; i++) ye[i] = T(0);
#pragma unroll
  for (int g = 0; g < 8; g++) {
    const T *__restrict__ G = Gt + 24 * g;
    T H[9];
#pragma unroll
    for (int i = 0; i < 9; i++) H[i] = T(0);
#pragma unroll
    for (int a = 0; a < 8; a++)
#pragma unroll
      for (int c = 0; c < 3; c++)
#pragma unroll
        for (int d = 0; d < 3; d++) H[3 * c + d] += xe[3 * a + c] * G[3 * a + d];
    const T s = sg[g];
    const T tr = lam * (H[0] + H[4] + H[8]);
    T S[9];
    S[0] = s * (tr + T(2) * mu * H[0]);
    S[4] = s * (tr + T(2) * mu * H[4]);
    S[8] = s * (tr + T(2) * mu * H[8]);
    S[1] = S[3] = s * mu * (H[1] + H[3]);
    S[2] = S[6] = s * mu * (H[2] + H[6]);
    S[5] = S[7] = s * mu * (H[5] + H[7]);
#pragma unroll
    for (int a = 0; a < 8; a++)
      if (mask & (1 << a)) {
#pragma unroll
        for (int c = 0; c < 3; c++)
          ye[3 * a + c] += S[3 * c] * G[3 * a] + S[3 * c + 1] * G[3 * a + 1] +
                           S[3 * c + 2] * G[3 * a + 2];
      }
  }
}

// LDS values per node of the vector-ALU tile apply.  Three: padding FP32
// nodes to 16 bytes (one 128-bit access per corner) costs more in tiles per
// CU than it saves in LDS instructions.
template <class T>
constexpr int tile_stride() {
  return 3;
}
// x slots: 3 values per node as well; 16-byte FP32 slots (one 128-bit read
// per corner) measured slower for the tiles per CU they cost
template <class T>
constexpr int tile_xstride() {
  return 3;
}

template <class T, int S>
__device__ __forceinline__ void lds_get3(const T *p, T &a, T &b, T &c) {
  if constexpr (S == 4 && sizeof(T) == 4) {
    const float4 v = *reinterpret_cast<const float4 *>(p);
    a = v.x;
    b = v.y;
    c = v.z;
  } else {
    a = p[0];
    b = p[1];
    c = p[2];
  }
}

// p[0..2] += s * (a, b, c), one FMA per value (FP32 vector-ALU accumulation)
__device__ __forceinline__ void lds_fma3(float *p, float s, float a, float b, float c) {
  p[0] = __builtin_fmaf(s, a, p[0]);
  p[1] = __builtin_fmaf(s, b, p[1]);
  p[2] = __builtin_fmaf(s, c, p[2]);
}

template <class T, int S>
__device__ __forceinline__ void lds_add3(T *p, T a, T b, T c) {
  if constexpr (S == 4 && sizeof(T) == 4) {
    float4 v = *reinterpret_cast<float4 *>(p);
    v.x += a;
    v.y += b;
    v.z += c;
    *reinterpret_cast<float4 *>(p) = v;
  } else if constexpr (sizeof(T) == 8) {
    // one writer per slot within a colour phase: a no-return LDS atomic add
    // is the plain read-add-write in one instruction (measured faster for
    // FP64; the FP32 LDS atomic is several times slower than read + write)
    __hip_atomic_fetch_add(p, a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    __hip_atomic_fetch_add(p + 1, b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    __hip_atomic_fetch_add(p + 2, c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  } else {
    p[0] += a;
    p[1] += b;
    p[2] += c;
  }
}

template <class T, int MODE, bool SINGLE_CLASS, bool NODAL, int NT>
__global__ __launch_bounds__(NT) void k_tile_apply(
    const int *__restrict__ node0, const int *__restrict__ ext_ptr, const int *__restrict__ ext,
    const int *__restrict__ nacc_dep, const int *__restrict__ el_ptr,
    const uint2 *__restrict__ el, int ncolors, const int *__restrict__ comb_ptr,
    const int *__restrict__ comb_slot, const float *__restrict__ comb_w,
    const int *__restrict__ dep_ptr, const int *__restrict__ dep_conn,
    const float *__restrict__ dep_w, const uint8_t *__restrict__ bcmask,
    const T *__restrict__ K0, const T *__restrict__ s_elem, const T *__restrict__ x,
    const T *__restrict__ b, T *__restrict__ y, T *__restrict__ x_out, T *__restrict__ d,
    const T *__restrict__ dinv, double ca, double cc, int cflags, double *__restrict__ partials,
    int64_t vs, const int *__restrict__ run) {
  if (run && *run == 0) return;
  constexpr int S = tile_stride<T>();
  constexpr int SX = tile_xstride<T>();
  // FP32 element-density products scale in the accumulation FMAs (24 fewer
  // multiplies per record); FP64 accumulates by LDS atomic adds, so it
  // scales inside the product
  constexpr bool FMA_ACC = sizeof(T) == 4 && !NODAL;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  __shared__ double red[NT / 64];
  const int t = xcd_block(blockIdx.x, gridDim.x);
  const int tid = threadIdx.x;
  {  // right-hand side blockIdx.y of a batch of vectors, stride vs
    const int64_t off = (int64_t)blockIdx.y * vs;
    x += off;
    if (b) b += off;
    if (y) y += off;
    if (x_out) x_out += off;
    if (d) d += off;
    if (partials) partials += (int64_t)blockIdx.y * gridDim.x;
  }
  const int n0 = node0[t], nown = node0[t + 1] - n0;
  const int e0 = ext_ptr[t], next = ext_ptr[t + 1] - e0;
  const int nslot = nown + next;
  const int nacc = nown + nacc_dep[t];
  T *xs = reinterpret_cast<T *>(smem_raw);
  T *ys = xs + SX * ((nslot + 7) & ~7);
  const int q0 = el_ptr[t], nel = el_ptr[t + 1] - q0;
  // the first round's records go out before the staging loads
  uint2 w0 = make_uint2(0, 0), w1 = w0, w2 = w0;
  if (tid < nel) {
    w0 = el[3 * (size_t)(q0 + tid)];
    w1 = el[3 * (size_t)(q0 + tid) + 1];
    w2 = el[3 * (size_t)(q0 + tid) + 2];
  }
  tile_stage<T, SX, NT, S>(n0, nown, ext + e0, next, nacc, dep_ptr, dep_conn, dep_w,
                                     bcmask, x, xs, ys);
  __syncthreads();
  for (int r0 = 0; r0 < nel; r0 += NT) {
    const int k = r0 + tid;
    T ye[24];
    T sc = T(0);
    int lc[8];
    int mask = 0, col = -1;
    if (k < nel) {
      lc[0] = w0.x & 0xffff; lc[1] = w0.x >> 16;
      lc[2] = w0.y & 0xffff; lc[3] = w0.y >> 16;
      lc[4] = w1.x & 0xffff; lc[5] = w1.x >> 16;
      lc[6] = w1.y & 0xffff; lc[7] = w1.y >> 16;
      const int elem = (int)w2.x;
      mask = w2.y & 0xff;
      col = (w2.y >> 8) & 0xff;
      const int cls = SINGLE_CLASS ? 0 : (int)((w2.y >> 16) & 0xff);
      T xe[24];
#pragma unroll
      for (int a = 0; a < 8; a++)
        lds_get3<T, SX>(xs + SX * lc[a], xe[3 * a], xe[3 * a + 1], xe[3 * a + 2]);
      if constexpr (NODAL) {
        nodal_elem<T>(K0 + NODAL_TABLE * cls, s_elem + 8 * (size_t)elem, xe, mask, ye);
      } else {
        // FP32: the record's element word holds the scale itself
        sc = FMA_ACC ? __uint_as_float(w2.x) : s_elem[elem];
        refl_elem<T, !FMA_ACC>(K0 + REFL_TABLE * cls, sc, xe, ye);
      }
    }
    // next round's records load while this round accumulates; the index is
    // clamped to the tile's last record (nel >= 1 inside the loop), so the
    // loads are unconditional and the waits for this round's scale need not
    // drain them
    {
      const size_t kn = 3 * (size_t)(q0 + min(k + NT, nel - 1));
      w0 = el[kn];
      w1 = el[kn + 1];
      w2 = el[kn + 2];
    }
    // elements are sorted by colour: this round only spans the colours of
    // its first and last element, and most waves hold a single colour (the
    // loads follow the round's arithmetic: issued earlier, they measured
    // slower)
    const int c_lo = (int)((el[3 * (size_t)(q0 + r0) + 2].y >> 8) & 0xff);
    const int c_hi = (int)((el[3 * (size_t)(q0 + min(r0 + NT, nel) - 1) + 2].y >> 8) & 0xff);
    for (int c = c_lo; c <= c_hi; c++) {
      auto acc = [&](int a) {
        T *p = ys + __umul24(lc[a], S);  // slots are below 2^16
        if constexpr (FMA_ACC)
          lds_fma3(p, sc, ye[3 * a], ye[3 * a + 1], ye[3 * a + 2]);
        else
          lds_add3<T, S>(p, ye[3 * a], ye[3 * a + 1], ye[3 * a + 2]);
      };
      if (col == c) {
#pragma unroll
        for (int a = 0; a < 8; a++)
          if (mask & (1 << a)) acc(a);
      }
      __syncthreads();
    }
  }
  tile_finish<T, MODE, S, NT, false, SX>(t, n0, nown, comb_ptr, comb_slot, comb_w,
                                                   bcmask, xs, ys, x, b, y, x_out, d, dinv, ca,
                                                   cc, cflags, partials, red, 0, nacc > nown);
}

// ---------------------------------------------------------------------------
// FP32 owner-tile apply on the matrix cores.  A wave takes 16 elements of one
// colour and class at a time and forms Y (24 x 16) = K0 (24 x 24) X (24 x 16)
// with v_mfma_f32_16x16x32_bf16.  The dof orderings are permuted so that lane
// group g = l / 16 of element n = l % 16 owns corners 2g and 2g + 1: K index
// 8g + 4c + j is component j of corner 2g + c, and row 4g + j of row block b
// is component j of corner 2g + b (j = 3 pads, zero in A).  LDS keeps 4
// floats per node, so a corner is one 16-byte read and one 16-byte update.
// Every FP32 operand is split into three bf16 parts (hi + mid + lo, 24
// significant bits) and the six products down to 2^-16 relative are summed in
// FP32: the product keeps FP32 accuracy, which the coarse-grid correction
// needs (residuals of smooth errors cancel to a small fraction of |K||x|).
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));
union Frag {
  uint4 u;
  bf16x8_t v;
};

// three bf16 parts of a pair of floats, packed two per word
__device__ __forceinline__ void split3(float a, float b, uint32_t &h, uint32_t &m, uint32_t &l) {
  union P {
    __bf16 e[2];
    uint32_t u;
  } ph, pm, pl;
  ph.e[0] = (__bf16)a;
  ph.e[1] = (__bf16)b;
  const float ra = a - (float)ph.e[0], rb = b - (float)ph.e[1];
  pm.e[0] = (__bf16)ra;
  pm.e[1] = (__bf16)rb;
  pl.e[0] = (__bf16)(ra - (float)pm.e[0]);
  pl.e[1] = (__bf16)(rb - (float)pm.e[1]);
  h = ph.u;
  m = pm.u;
  l = pl.u;
}

// Every wave owns a contiguous share of the tile's nodes (and the hanging
// nodes whose first owned parent it owns) and works through the batches of
// the elements touching them, in order and with the rows of its own corners
// only: no two waves write the same accumulator, so the element phase needs
// no barriers, and records are fetched three batches ahead and x one batch
// ahead of the batch on the matrix cores.
// stage x as three bf16 planes (hi, mid, lo; 4 bf16 per node: x, y, z, 0),
// each np nodes long: the matrix-core operands are read ready-split
__device__ __forceinline__ void put_planes(uint2 *ph, int np, int pos, float vx, float vy,
                                           float vz) {
  uint32_t h0, m0, l0, h1, m1, l1;
  split3(vx, vy, h0, m0, l0);
  split3(vz, 0.0f, h1, m1, l1);
  ph[pos] = make_uint2(h0, h1);
  ph[pos + np] = make_uint2(m0, m1);
  ph[pos + 2 * np] = make_uint2(l0, l1);
}

__device__ __forceinline__ void tile_stage_planes(int n0, int nown, const int *__restrict__ ext_l,
                                                  int next, int nacc, int np,
                                                  const int *__restrict__ dep_ptr,
                                                  const int *__restrict__ dep_conn,
                                                  const float *__restrict__ dep_w,
                                                  const uint8_t *__restrict__ bcmask,
                                                  const float *__restrict__ x, uint2 *ph,
                                                  float *ys) {
  const int tid = threadIdx.x;
  for (int base = 0; base < nown; base += 2 * TILE_THREADS) {  // two nodes in flight
    float v[2][3];
    uint8_t mk[2];
#pragma unroll
    for (int k = 0; k < 2; k++) {
      const int j = min(base + k * TILE_THREADS + tid, nown - 1);
      mk[k] = bcmask[n0 + j];
      v[k][0] = x[3 * (size_t)(n0 + j)];
      v[k][1] = x[3 * (size_t)(n0 + j) + 1];
      v[k][2] = x[3 * (size_t)(n0 + j) + 2];
    }
#pragma unroll
    for (int k = 0; k < 2; k++) {
      const int j = base + k * TILE_THREADS + tid;
      if (j < nown)
        put_planes(ph, np, tile_pos(j), (mk[k] & 1) ? 0.0f : v[k][0], (mk[k] & 2) ? 0.0f : v[k][1],
                   (mk[k] & 4) ? 0.0f : v[k][2]);
    }
  }
  for (int i = tid; i < next; i += TILE_THREADS) {
    float v[3];
    load_node<float>(ext_l[i], x, dep_ptr, dep_conn, dep_w, bcmask, v);
    put_planes(ph, np, tile_pos(nown + i), v[0], v[1], v[2]);
  }
  for (int i = tid; i < 4 * ((nacc + 7) & ~7); i += TILE_THREADS) ys[i] = 0.0f;
}

template <int MODE, bool SINGLE_CLASS>
__global__ __launch_bounds__(TILE_THREADS) void k_tile_apply_mfma(
    const int *__restrict__ node0, const int *__restrict__ ext_ptr, const int *__restrict__ ext,
    const int *__restrict__ nacc_dep, const int *__restrict__ mb_ptr,
    const uint32_t *__restrict__ mel, int ncolors, const uint4 *__restrict__ afrag,
    const int *__restrict__ comb_ptr, const int *__restrict__ comb_slot,
    const float *__restrict__ comb_w, const int *__restrict__ dep_ptr,
    const int *__restrict__ dep_conn, const float *__restrict__ dep_w,
    const uint8_t *__restrict__ bcmask, const float *__restrict__ s_elem,
    const float *__restrict__ x, const float *__restrict__ b, float *__restrict__ y,
    float *__restrict__ x_out, float *__restrict__ d, const float *__restrict__ dinv, double ca,
    double cc, int cflags, double *__restrict__ partials, int64_t vs,
    const int *__restrict__ run) {
  if (run && *run == 0) return;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  __shared__ double red[TILE_THREADS / 64];
  const int t = xcd_block(blockIdx.x, gridDim.x);
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  {  // right-hand side blockIdx.y of a batch of vectors, stride vs
    const int64_t off = (int64_t)blockIdx.y * vs;
    x += off;
    if (b) b += off;
    if (y) y += off;
    if (x_out) x_out += off;
    if (d) d += off;
    if (partials) partials += (int64_t)blockIdx.y * gridDim.x;
  }
  const int en = lane & 15, g = lane >> 4;
  const int n0 = node0[t], nown = node0[t + 1] - n0;
  const int e0 = ext_ptr[t], next = ext_ptr[t + 1] - e0;
  const int nslot = nown + next;
  const int nacc = nown + nacc_dep[t];
  // LDS: x as three bf16 planes of np nodes (8 bytes per node each), then
  // the FP32 accumulators (16 bytes per node)
  const int np = (nslot + 7) & ~7;
  uint2 *ph = reinterpret_cast<uint2 *>(smem_raw);
  float *ys = reinterpret_cast<float *>(ph + 3 * np);
  float4 *ys4 = reinterpret_cast<float4 *>(ys);
  const int *cp = mb_ptr + (size_t)t * (MFMA_WAVES + 1);
  // wave-uniform batch range: keeps the loop control scalar
  const int bb0 = __builtin_amdgcn_readfirstlane(cp[wv]);
  const int bb1 = __builtin_amdgcn_readfirstlane(cp[wv + 1]);
  // record words: 0-3 corner positions (two per word), 4 FP32 stiffness
  // scale, 5 mask | colour << 8 | class << 16
  struct Rec {
    uint32_t p, m;
    float s;
  };
  auto load_rec = [&](int bt, Rec &r) {
    const uint32_t *rec = mel + 6 * ((size_t)bt * 16 + en);
    r.p = rec[g];
    r.s = __uint_as_float(rec[4]);
    r.m = rec[5];
  };
  Rec r0{0u, 0u, 0.0f}, r1 = r0, r2 = r0, r3 = r0;
  if (bb0 < bb1) load_rec(bb0, r0);
  if (bb0 + 1 < bb1) load_rec(bb0 + 1, r1);
  if (bb0 + 2 < bb1) load_rec(bb0 + 2, r2);
  tile_stage_planes(n0, nown, ext + e0, next, nacc, np, dep_ptr, dep_conn, dep_w, bcmask, x, ph, ys);
  // A fragments of K0: [class][part hi, mid, lo][row block][lane]
  Frag a[3][2];
  if (SINGLE_CLASS) {
#pragma unroll
    for (int p = 0; p < 3; p++) {
      a[p][0].u = afrag[(2 * p) * 64 + lane];
      a[p][1].u = afrag[(2 * p + 1) * 64 + lane];
    }
    // keep them in registers: without this the compiler reloads them from
    // memory in every batch to save registers
#pragma unroll
    for (int p = 0; p < 3; p++)
#pragma unroll
      for (int q = 0; q < 2; q++)
        asm volatile("" : "+v"(a[p][q].u.x), "+v"(a[p][q].u.y), "+v"(a[p][q].u.z),
                     "+v"(a[p][q].u.w));
  }
  __syncthreads();
  // B fragments of a batch: the planes of its two corners per lane
  auto load_b = [&](uint32_t pr, Frag *B) {
    const int q0 = pr & 0xffff, q1 = pr >> 16;
    const uint2 a0 = ph[q0], a1 = ph[q1], b0 = ph[q0 + np], b1 = ph[q1 + np];
    const uint2 c0 = ph[q0 + 2 * np], c1 = ph[q1 + 2 * np];
    B[0].u = make_uint4(a0.x, a0.y, a1.x, a1.y);
    B[1].u = make_uint4(b0.x, b0.y, b1.x, b1.y);
    B[2].u = make_uint4(c0.x, c0.y, c1.x, c1.y);
  };
  auto compute = [&](const Rec &cur, const Frag *B) {
    if (!SINGLE_CLASS) {
      const int cls = __builtin_amdgcn_readfirstlane((int)((cur.m >> 16) & 0xff));
      const uint4 *af = afrag + (size_t)cls * 384;
#pragma unroll
      for (int p = 0; p < 3; p++) {
        a[p][0].u = af[(2 * p) * 64 + lane];
        a[p][1].u = af[(2 * p + 1) * 64 + lane];
      }
    }
    // one accumulation chain per row block, small terms first; the two
    // chains interleave on the matrix core
    f32x4_t acc[2];
    const f32x4_t zero = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int blk = 0; blk < 2; blk++)
      acc[blk] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[2][blk].v, B[0].v, zero, 0, 0, 0);
#pragma unroll
    for (int blk = 0; blk < 2; blk++)
      acc[blk] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[1][blk].v, B[1].v, acc[blk], 0, 0, 0);
#pragma unroll
    for (int blk = 0; blk < 2; blk++)
      acc[blk] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0][blk].v, B[2].v, acc[blk], 0, 0, 0);
#pragma unroll
    for (int blk = 0; blk < 2; blk++)
      acc[blk] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[1][blk].v, B[0].v, acc[blk], 0, 0, 0);
#pragma unroll
    for (int blk = 0; blk < 2; blk++)
      acc[blk] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0][blk].v, B[1].v, acc[blk], 0, 0, 0);
#pragma unroll
    for (int blk = 0; blk < 2; blk++)
      acc[blk] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0][blk].v, B[0].v, acc[blk], 0, 0, 0);
    const int mask = cur.m & 0xff;
    // rows of block b are the components of corner 2g + b
    if (mask & (1 << (2 * g))) {
      float4 w = ys4[cur.p & 0xffff];
      w.x += cur.s * acc[0][0];
      w.y += cur.s * acc[0][1];
      w.z += cur.s * acc[0][2];
      ys4[cur.p & 0xffff] = w;
    }
    if (mask & (2 << (2 * g))) {
      float4 w = ys4[cur.p >> 16];
      w.x += cur.s * acc[1][0];
      w.y += cur.s * acc[1][1];
      w.z += cur.s * acc[1][2];
      ys4[cur.p >> 16] = w;
    }
  };
  // ring of four records (fetched three batches ahead) and two sets of B
  // fragments (read one batch ahead), unrolled so that nothing is copied
  Frag B0[3], B1[3];
  if (bb0 < bb1) load_b(r0.p, B0);
  auto step = [&](int bt, const Rec &cur, const Rec &nxt, Rec &fill, const Frag *Bc, Frag *Bn) {
    if (bt + 3 < bb1) load_rec(bt + 3, fill);
    if (bt + 1 < bb1) load_b(nxt.p, Bn);  // x is read-only during this phase
    compute(cur, Bc);
  };
  int bt = bb0;
  for (; bt + 4 <= bb1; bt += 4) {
    step(bt, r0, r1, r3, B0, B1);
    step(bt + 1, r1, r2, r0, B1, B0);
    step(bt + 2, r2, r3, r1, B0, B1);
    step(bt + 3, r3, r0, r2, B1, B0);
  }
  if (bt < bb1) step(bt, r0, r1, r3, B0, B1);
  if (bt + 1 < bb1) step(bt + 1, r1, r2, r0, B1, B0);
  if (bt + 2 < bb1) step(bt + 2, r2, r3, r1, B0, B1);
  __syncthreads();
  tile_finish<float, MODE, 4, TILE_THREADS, true>(
      t, n0, nown, comb_ptr, comb_slot, comb_w, bcmask, reinterpret_cast<const float *>(ph), ys, x,
      b, y, x_out, d, dinv, ca, cc, cflags, partials, red, np);
}

// ---------------------------------------------------------------------------
// FP32 owner-tile apply on the matrix cores with FP32 inputs
// (v_mfma_f32_16x16x4_f32): x staged once, 3 floats per slot, and K0 as FP32
// A fragments; Y (32 x 16) = K0 (32 x 24) X (24 x 16) in six K-steps of 4
// per row block.  K index d = 4 s + g (step s, lane group g = lane / 16) is
// dof d of element n = lane % 16 (corner d / 3, component d % 3); output row
// 4 a + j is component j of corner a (j = 3 pads: zero rows of A), so row
// block b leaves lane group g the three components of corner 4 b + g.
// Batches, wave ownership of the accumulators, record ring and epilogues are
// those of k_tile_apply_mfma; products and sums are FP32 throughout.
template <int MODE, bool SINGLE_CLASS>
__global__ __launch_bounds__(TILE_THREADS) void k_tile_apply_mfma32(
    const int *__restrict__ node0, const int *__restrict__ ext_ptr, const int *__restrict__ ext,
    const int *__restrict__ nacc_dep, const int *__restrict__ mb_ptr,
    const uint32_t *__restrict__ mel, const float *__restrict__ afrag32,
    const int *__restrict__ comb_ptr, const int *__restrict__ comb_slot,
    const float *__restrict__ comb_w, const int *__restrict__ dep_ptr,
    const int *__restrict__ dep_conn, const float *__restrict__ dep_w,
    const uint8_t *__restrict__ bcmask, const float *__restrict__ x, const float *__restrict__ b,
    float *__restrict__ y, float *__restrict__ x_out, float *__restrict__ d,
    const float *__restrict__ dinv, double ca, double cc, int cflags,
    double *__restrict__ partials, int64_t vs, const int *__restrict__ run) {
  if (run && *run == 0) return;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  __shared__ double red[TILE_THREADS / 64];
  const int t = xcd_block(blockIdx.x, gridDim.x);
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  {  // right-hand side blockIdx.y of a batch of vectors, stride vs
    const int64_t off = (int64_t)blockIdx.y * vs;
    x += off;
    if (b) b += off;
    if (y) y += off;
    if (x_out) x_out += off;
    if (d) d += off;
    if (partials) partials += (int64_t)blockIdx.y * gridDim.x;
  }
  const int en = lane & 15, g = lane >> 4;
  const int n0 = node0[t], nown = node0[t + 1] - n0;
  const int e0 = ext_ptr[t], next = ext_ptr[t + 1] - e0;
  const int nslot = nown + next;
  const int nacc = nown + nacc_dep[t];
  float *xs = reinterpret_cast<float *>(smem_raw);
  float *ys = xs + 3 * ((nslot + 7) & ~7);
  float4 *ys4 = reinterpret_cast<float4 *>(ys);
  const int *cp = mb_ptr + (size_t)t * (MFMA_WAVES + 1);
  const int bb0 = __builtin_amdgcn_readfirstlane(cp[wv]);
  const int bb1 = __builtin_amdgcn_readfirstlane(cp[wv + 1]);
  // record words: 0-3 corner positions (two per word), 4 FP32 stiffness
  // scale, 5 mask | colour << 8 | class << 16
  struct Rec {
    uint32_t p0, p1, p2, p3, m;
    float s;
  };
  auto load_rec = [&](int bt, Rec &r) {
    const uint2 *rec = reinterpret_cast<const uint2 *>(mel + 6 * ((size_t)bt * 16 + en));
    const uint2 a = rec[0], c = rec[1], e = rec[2];
    r.p0 = a.x;
    r.p1 = a.y;
    r.p2 = c.x;
    r.p3 = c.y;
    r.s = __uint_as_float(e.x);
    r.m = e.y;
  };
  Rec r0{0u, 0u, 0u, 0u, 0u, 0.0f}, r1 = r0, r2 = r0, r3 = r0;
  if (bb0 < bb1) load_rec(bb0, r0);
  if (bb0 + 1 < bb1) load_rec(bb0 + 1, r1);
  if (bb0 + 2 < bb1) load_rec(bb0 + 2, r2);
  tile_stage<float, 3, TILE_THREADS, 4>(n0, nown, ext + e0, next, nacc, dep_ptr, dep_conn, dep_w,
                                        bcmask, x, xs, ys);
  // A fragments: [class][row block][step][lane]
  float a[2][6];
  if (SINGLE_CLASS) {
#pragma unroll
    for (int blk = 0; blk < 2; blk++)
#pragma unroll
      for (int s = 0; s < 6; s++) a[blk][s] = afrag32[(blk * 6 + s) * 64 + lane];
#pragma unroll
    for (int blk = 0; blk < 2; blk++)
#pragma unroll
      for (int s = 0; s < 6; s++) asm volatile("" : "+v"(a[blk][s]));
  }
  // dof 4 s + g of an element: corner c_s or c_s + 1 (up), component comp
  // (constants of the lane group)
  bool up[6];
  int comp[6];
#pragma unroll
  for (int s = 0; s < 6; s++) {
    const int dd = 4 * s + g, c0 = (4 * s) / 3;
    up[s] = dd >= 3 * (c0 + 1);
    comp[s] = dd - 3 * (up[s] ? c0 + 1 : c0);
  }
  __syncthreads();
  // corner slot a of a record (compile-time a)
  auto slot_of = [](const Rec &r, int c) -> int {
    const uint32_t w = c < 2 ? r.p0 : c < 4 ? r.p1 : c < 6 ? r.p2 : r.p3;
    return (int)((c & 1) ? (w >> 16) : (w & 0xffff));
  };
  auto load_b = [&](const Rec &r, float *B) {
#pragma unroll
    for (int s = 0; s < 6; s++) {
      const int c0 = (4 * s) / 3;
      const int sl = up[s] ? slot_of(r, c0 + 1) : slot_of(r, c0);
      B[s] = xs[3 * sl + comp[s]];
    }
  };
  auto compute = [&](const Rec &cur, const float *B) {
    if (!SINGLE_CLASS) {
      const int cls = __builtin_amdgcn_readfirstlane((int)((cur.m >> 16) & 0xff));
      const float *af = afrag32 + (size_t)cls * 768;
#pragma unroll
      for (int blk = 0; blk < 2; blk++)
#pragma unroll
        for (int s = 0; s < 6; s++) a[blk][s] = af[(blk * 6 + s) * 64 + lane];
    }
    f32x4_t acc0 = {0.0f, 0.0f, 0.0f, 0.0f}, acc1 = acc0;
#pragma unroll
    for (int s = 0; s < 6; s++) {
      acc0 = __builtin_amdgcn_mfma_f32_16x16x4f32(a[0][s], B[s], acc0, 0, 0, 0);
      acc1 = __builtin_amdgcn_mfma_f32_16x16x4f32(a[1][s], B[s], acc1, 0, 0, 0);
    }
    const int mask = cur.m & 0xff;
    // corner g (row block 0) and corner 4 + g (row block 1)
    const uint32_t wlo = (g & 2) ? cur.p1 : cur.p0, whi = (g & 2) ? cur.p3 : cur.p2;
    const int slo = (int)((g & 1) ? (wlo >> 16) : (wlo & 0xffff));
    const int shi = (int)((g & 1) ? (whi >> 16) : (whi & 0xffff));
    if (mask & (1 << g)) {
      float4 w = ys4[slo];
      w.x = __builtin_fmaf(cur.s, acc0[0], w.x);
      w.y = __builtin_fmaf(cur.s, acc0[1], w.y);
      w.z = __builtin_fmaf(cur.s, acc0[2], w.z);
      ys4[slo] = w;
    }
    if (mask & (16 << g)) {
      float4 w = ys4[shi];
      w.x = __builtin_fmaf(cur.s, acc1[0], w.x);
      w.y = __builtin_fmaf(cur.s, acc1[1], w.y);
      w.z = __builtin_fmaf(cur.s, acc1[2], w.z);
      ys4[shi] = w;
    }
  };
  // ring of four records (fetched three batches ahead) and two sets of B
  // operands (read one batch ahead)
  float B0[6], B1[6];
  if (bb0 < bb1) load_b(r0, B0);
  auto step = [&](int bt, const Rec &cur, const Rec &nxt, Rec &fill, const float *Bc, float *Bn) {
    if (bt + 3 < bb1) load_rec(bt + 3, fill);
    if (bt + 1 < bb1) load_b(nxt, Bn);  // x is read-only during this phase
    compute(cur, Bc);
  };
  int bt = bb0;
  for (; bt + 4 <= bb1; bt += 4) {
    step(bt, r0, r1, r3, B0, B1);
    step(bt + 1, r1, r2, r0, B1, B0);
    step(bt + 2, r2, r3, r1, B0, B1);
    step(bt + 3, r3, r0, r2, B1, B0);
  }
  if (bt < bb1) step(bt, r0, r1, r3, B0, B1);
  if (bt + 1 < bb1) step(bt + 1, r1, r2, r0, B1, B0);
  if (bt + 2 < bb1) step(bt + 2, r2, r3, r1, B0, B1);
  __syncthreads();
  tile_finish<float, MODE, 4, TILE_THREADS, false, 3>(t, n0, nown, comb_ptr, comb_slot, comb_w,
                                                      bcmask, xs, ys, x, b, y, x_out, d, dinv, ca,
                                                      cc, cflags, partials, red);
}

template <class T>
void launch_tile_apply(const Tiles &tl, const T *K0, const T *s_elem, const T *x, int mode,
                       const AsmArgs &g, hipStream_t st, int nrhs, int64_t vs) {
  if (tl.n_tiles == 0) return;
  const dim3 grid(tl.n_tiles, nrhs);
  const size_t sh = ((size_t)tile_xstride<T>() * round8(tl.max_slots) +
                     (size_t)tile_stride<T>() * round8(tl.max_acc)) * sizeof(T);
  const T *b = (const T *)g.b;
  T *y = (T *)g.y, *xo = (T *)g.x_out, *d = (T *)g.d;
  const T *dinv = (const T *)g.dinv;
  const bool single = tl.single_class;
  if constexpr (sizeof(T) == 4) {
    if (tl.use_mfma && tl.mfma_f32 && !tl.nodal) {
      const size_t sh3 = 12 * (size_t)round8(tl.max_slots) + 16 * (size_t)round8(tl.max_acc);
#define TMRX_MTILE32(M, S)                                                                      \
  hipLaunchKernelGGL((k_tile_apply_mfma32<M, S>), grid, dim3(TILE_THREADS), sh3, st, tl.node0,  \
                     tl.ext_ptr, tl.ext, tl.nacc_dep, tl.mb_ptr,                                \
                     reinterpret_cast<const uint32_t *>(tl.mel), tl.afrag32, tl.comb_ptr,       \
                     tl.comb_slot, tl.comb_w, tl.dep_ptr, tl.dep_conn, tl.dep_w, tl.bcmask,     \
                     (const float *)x, (const float *)b, (float *)y, (float *)xo, (float *)d,    \
                     (const float *)dinv, g.ca, g.cc, g.cflags, g.partials, vs, tl.run)
#define TMRX_MTILE32B(M)    \
  if (single) {             \
    TMRX_MTILE32(M, true);  \
  } else {                  \
    TMRX_MTILE32(M, false); \
  }
      switch (mode) {
        case ASM_STORE: TMRX_MTILE32B(ASM_STORE); break;
        case ASM_RESID: TMRX_MTILE32B(ASM_RESID); break;
        case ASM_CHEB: TMRX_MTILE32B(ASM_CHEB); break;
        case ASM_DOT: TMRX_MTILE32B(ASM_DOT); break;
      }
#undef TMRX_MTILE32B
#undef TMRX_MTILE32
      TMRX_LAUNCH_CHECK();
      return;
    }
    if (tl.use_mfma && !tl.nodal) {
      const size_t sh4 = 24 * (size_t)round8(tl.max_slots) + 16 * (size_t)round8(tl.max_acc);
#define TMRX_MTILE(M, S)                                                                         \
  hipLaunchKernelGGL((k_tile_apply_mfma<M, S>), grid, dim3(TILE_THREADS), sh4, st,               \
                     tl.node0, tl.ext_ptr, tl.ext, tl.nacc_dep, tl.mb_ptr,                       \
                     reinterpret_cast<const uint32_t *>(tl.mel), tl.ncolors, tl.afrag, tl.comb_ptr, \
                     tl.comb_slot, tl.comb_w, tl.dep_ptr, tl.dep_conn, tl.dep_w, tl.bcmask,      \
                     (const float *)s_elem, (const float *)x, (const float *)b, (float *)y,      \
                     (float *)xo, (float *)d, (const float *)dinv, g.ca, g.cc, g.cflags,          \
                     g.partials, vs, tl.run)
#define TMRX_MTILE2(M)    \
  if (single) {           \
    TMRX_MTILE(M, true);  \
  } else {                \
    TMRX_MTILE(M, false); \
  }
      switch (mode) {
        case ASM_STORE: TMRX_MTILE2(ASM_STORE); break;
        case ASM_RESID: TMRX_MTILE2(ASM_RESID); break;
        case ASM_CHEB: TMRX_MTILE2(ASM_CHEB); break;
        case ASM_DOT: TMRX_MTILE2(ASM_DOT); break;
      }
#undef TMRX_MTILE2
#undef TMRX_MTILE
      TMRX_LAUNCH_CHECK();
      return;
    }
  }
  // FP32 element-density records carry the scale in their element word
  const TileElem *recs = (sizeof(T) == 4 && !tl.nodal) ? tl.el32 : tl.el;
  if (!recs) {  // tiles not built: report like a failed launch
    note_launch(hipErrorInvalidValue);
    return;
  }
#define TMRX_TILE_NT(M, S, N, NTH)                                                              \
  hipLaunchKernelGGL((k_tile_apply<T, M, S, N, NTH>), grid, dim3(NTH), sh, st, tl.node0,          \
                     tl.ext_ptr, tl.ext, tl.nacc_dep, tl.el_ptr,                                  \
                     reinterpret_cast<const uint2 *>(recs), tl.ncolors, tl.comb_ptr,              \
                     tl.comb_slot, tl.comb_w, tl.dep_ptr, tl.dep_conn, tl.dep_w, tl.bcmask, K0,   \
                     s_elem, x, b, y, xo, d, dinv, g.ca, g.cc, g.cflags, g.partials, vs, tl.run)
#define TMRX_TILE(M, S)                    \
  if (tl.threads == 512) {                 \
    if (tl.nodal)                          \
      TMRX_TILE_NT(M, S, true, 512);       \
    else                                   \
      TMRX_TILE_NT(M, S, false, 512);      \
  } else {                                 \
    if (tl.nodal)                          \
      TMRX_TILE_NT(M, S, true, 256);       \
    else                                   \
      TMRX_TILE_NT(M, S, false, 256);      \
  }
#define TMRX_TILE2(M) \
  if (single) {       \
    TMRX_TILE(M, true); \
  } else {            \
    TMRX_TILE(M, false); \
  }
  switch (mode) {
    case ASM_STORE: TMRX_TILE2(ASM_STORE); break;
    case ASM_RESID: TMRX_TILE2(ASM_RESID); break;
    case ASM_CHEB: TMRX_TILE2(ASM_CHEB); break;
    case ASM_DOT: TMRX_TILE2(ASM_DOT); break;
  }
#undef TMRX_TILE2
#undef TMRX_TILE
#undef TMRX_TILE_NT
  TMRX_LAUNCH_CHECK();
}

// ---------------------------------------------------------------------------
// p-norm von Mises functional and its derivatives (FP64)
__device__ __forceinline__ void vm_terms(const double s[6], double ys, double *J, double dJ[6]) {
  const double a = s[0] - s[1], b = s[1] - s[2], c = s[0] - s[2];
  *J = 0.5 * (a * a + b * b + c * c) + 3.0 * (s[3] * s[3] + s[4] * s[4] + s[5] * s[5]);
  dJ[0] = 2.0 * s[0] - s[1] - s[2];
  dJ[1] = 2.0 * s[1] - s[0] - s[2];
  dJ[2] = 2.0 * s[2] - s[0] - s[1];
  dJ[3] = 6.0 * s[3];
  dJ[4] = 6.0 * s[4];
  dJ[5] = 6.0 * s[5];
  (void)ys;
}

__global__ __launch_bounds__(256) void k_stress_sum(
    int ne, const int *__restrict__ conn, const int *__restrict__ dep_ptr,
    const int *__restrict__ dep_conn, const float *__restrict__ dep_w,
    const uint8_t *__restrict__ bcmask, const int *__restrict__ cls,
    const int16_t *__restrict__ lev, const double *__restrict__ rho,
    const double *__restrict__ rho_gp, const double *__restrict__ u, double P, double eps,
    double ys, const double *__restrict__ CB, const double *__restrict__ vol,
    double *__restrict__ partials) {
  __shared__ double red[4];
  const int e = blockIdx.x * 256 + threadIdx.x;
  double acc = 0.0;
  if (e < ne) {
    double ue[24];
    for (int a = 0; a < 8; a++) {
      double v[3];
      load_node<double>(conn[8 * e + a], u, dep_ptr, dep_conn, dep_w, bcmask, v);
      ue[3 * a] = v[0];
      ue[3 * a + 1] = v[1];
      ue[3 * a + 2] = v[2];
    }
    const int c = cls[e], l = lev[e];
    const double hs = ldexp(1.0, l);  // B scales as 1/h
    const double Ve = ldexp(vol[c], -3 * l);
    const double re = rho_gp ? 0.0 : rho[e];
    const double *G = CB + 8 * 144 * (size_t)c;
    for (int g = 0; g < 8; g++) {
      const double rh = rho_gp ? rho_gp[8 * (size_t)e + g] : re;
      const double r = rh / (eps * (1.0 - rh) + rh);
      double sg[6];
      for (int i = 0; i < 6; i++) {
        double t = 0.0;
        for (int j = 0; j < 24; j++) t += G[144 * g + 24 * i + j] * ue[j];
        sg[i] = hs * t;
      }
      double J, dJ[6];
      vm_terms(sg, ys, &J, dJ);
      const double fail = r * sqrt(J) / ys;
      acc += 0.125 * Ve * pow(fail, P);
    }
  }
  const double sum = block_sum(acc, red);
  if (threadIdx.x == 0) partials[blockIdx.x] = sum;
}

void launch_stress_sum(int nelem, const int *conn, const int *dep_ptr, const int *dep_conn,
                       const float *dep_w, const uint8_t *bcmask, const int *cls,
                       const int16_t *lev, const double *rho, const double *rho_gp,
                       const double *u, const StressArgs &sa, double *partials, hipStream_t st) {
  if (nelem == 0) return;
  hipLaunchKernelGGL(k_stress_sum, dim3((nelem + 255) / 256), dim3(256), 0, st, nelem, conn,
                     dep_ptr, dep_conn, dep_w, bcmask, cls, lev, rho, rho_gp, u, sa.P, sa.eps,
                     sa.ys, sa.CB, sa.vol, partials);
  TMRX_LAUNCH_CHECK();
}

// dfdu: per element g_e = sum_g coef (V/8) r^P vm^(P-2) / (2 ys^2) dJ^T (C0 B_g),
// staged per batch like the apply; explicit dfdrho_e = coef (V/8) r^(P-1) r'
// sum_g vm^P.  vm = sqrt(J) / ys.
__global__ __launch_bounds__(WG_ELEMS) void k_stress_grad(
    const int *__restrict__ wg_node_ptr, const int *__restrict__ wg_nodes,
    const int *__restrict__ wg_elem_ptr, const uint4 *__restrict__ lconn,
    const uint8_t *__restrict__ color, int ncolors, const int *__restrict__ dep_ptr,
    const int *__restrict__ dep_conn, const float *__restrict__ dep_w,
    const uint8_t *__restrict__ bcmask, const int16_t *__restrict__ lev,
    const int *__restrict__ cls, const double *__restrict__ rho,
    const double *__restrict__ rho_gp, const double *__restrict__ u, double P, double eps,
    double ys, const double *__restrict__ CB, const double *__restrict__ vol,
    const double *__restrict__ sc, double *__restrict__ partial, double *__restrict__ dfdrho) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int wg = blockIdx.x, tid = threadIdx.x;
  const int nb = wg_node_ptr[wg], nn = wg_node_ptr[wg + 1] - nb;
  double *xs = reinterpret_cast<double *>(smem_raw);
  double *ys_acc = xs + 3 * nn;
  for (int i = tid; i < nn; i += WG_ELEMS) {
    double v[3];
    load_node<double>(wg_nodes[nb + i], u, dep_ptr, dep_conn, dep_w, bcmask, v);
    xs[3 * i] = v[0];
    xs[3 * i + 1] = v[1];
    xs[3 * i + 2] = v[2];
    ys_acc[3 * i] = ys_acc[3 * i + 1] = ys_acc[3 * i + 2] = 0.0;
  }
  __syncthreads();
  const int e0 = wg_elem_ptr[wg], ne = wg_elem_ptr[wg + 1] - e0;
  const double coef = sc[SC_ZQ];
  double ge[24];
  int lc[8];
  int col = -1;
  if (tid < ne) {
    const int e = e0 + tid;
    const uint4 q = lconn[e];
    lc[0] = q.x & 0xffff; lc[1] = q.x >> 16;
    lc[2] = q.y & 0xffff; lc[3] = q.y >> 16;
    lc[4] = q.z & 0xffff; lc[5] = q.z >> 16;
    lc[6] = q.w & 0xffff; lc[7] = q.w >> 16;
    col = color[e];
    double ue[24];
    for (int a = 0; a < 8; a++)
      for (int d = 0; d < 3; d++) ue[3 * a + d] = xs[3 * lc[a] + d];
    for (int j = 0; j < 24; j++) ge[j] = 0.0;
    const int c = cls[e], l = lev[e];
    const double hs = ldexp(1.0, l);
    const double Ve = ldexp(vol[c], -3 * l);
    const double re = rho_gp ? 0.0 : rho[e];
    const double *G = CB + 8 * 144 * (size_t)c;
    double sum_vmP = 0.0;
    for (int g = 0; g < 8; g++) {
      const double rh = rho_gp ? rho_gp[8 * (size_t)e + g] : re;
      const double den = eps * (1.0 - rh) + rh;
      const double r = rh / den, dr = eps / (den * den);
      double sg[6];
      for (int i = 0; i < 6; i++) {
        double t = 0.0;
        for (int j = 0; j < 24; j++) t += G[144 * g + 24 * i + j] * ue[j];
        sg[i] = hs * t;
      }
      double J, dJ[6];
      vm_terms(sg, ys, &J, dJ);
      const double vm = sqrt(J) / ys;
      sum_vmP += pow(vm, P);
      if (rho_gp) dfdrho[8 * (size_t)e + g] = coef * 0.125 * Ve * pow(r, P - 1.0) * dr * pow(vm, P);
      // coef (V/8) r^P vm^(P-2) / (2 ys^2) * dJ^T CB_g (with the 1/h factor)
      const double w = coef * 0.125 * Ve * pow(r, P) * pow(vm, P - 2.0) / (2.0 * ys * ys) * hs;
      for (int i = 0; i < 6; i++) {
        const double wi = w * dJ[i];
        for (int j = 0; j < 24; j++) ge[j] += wi * G[144 * g + 24 * i + j];
      }
    }
    if (!rho_gp) {
      const double den = eps * (1.0 - re) + re;
      const double r = re / den, dr = eps / (den * den);
      dfdrho[e] = coef * 0.125 * Ve * pow(r, P - 1.0) * dr * sum_vmP;
    }
  }
  for (int c = 0; c < ncolors; c++) {
    if (col == c) {
      for (int a = 0; a < 8; a++)
        for (int d = 0; d < 3; d++) ys_acc[3 * lc[a] + d] += ge[3 * a + d];
    }
    __syncthreads();
  }
  for (int i = tid; i < 3 * nn; i += WG_ELEMS) partial[3 * nb + i] = ys_acc[i];
}

void launch_stress_grad(const ElemBatches &eb, const int16_t *lev, const int *cls,
                        const double *rho, const double *rho_gp, const double *u,
                        const StressArgs &sa, const double *sc, double *partial,
                        double *dfdrho, hipStream_t st) {
  if (eb.n_wg == 0) return;
  const size_t sh = 2 * 3 * (size_t)eb.max_wg_nodes * sizeof(double);
  hipLaunchKernelGGL(k_stress_grad, dim3(eb.n_wg), dim3(WG_ELEMS), sh, st, eb.wg_node_ptr,
                     eb.wg_nodes, eb.wg_elem_ptr, reinterpret_cast<const uint4 *>(eb.lconn),
                     eb.color, eb.ncolors, eb.dep_ptr, eb.dep_conn, eb.dep_w, eb.bcmask, lev,
                     cls, rho, rho_gp, u, sa.P, sa.eps, sa.ys, sa.CB, sa.vol, sc, partial,
                     dfdrho);
  TMRX_LAUNCH_CHECK();
}

__global__ void k_elem_energy2(int ne, const int *__restrict__ conn, const int *__restrict__ dep_ptr,
                               const int *__restrict__ dep_conn, const float *__restrict__ dep_w,
                               const uint8_t *__restrict__ bcmask, const int *__restrict__ cls,
                               const double *__restrict__ K0, const double *__restrict__ u,
                               const double *__restrict__ v, double *__restrict__ out) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ne) return;
  double ue[24], ve[24];
  for (int a = 0; a < 8; a++) {
    double t[3];
    load_node<double>(conn[8 * e + a], u, dep_ptr, dep_conn, dep_w, bcmask, t);
    ue[3 * a] = t[0];
    ue[3 * a + 1] = t[1];
    ue[3 * a + 2] = t[2];
    load_node<double>(conn[8 * e + a], v, dep_ptr, dep_conn, dep_w, bcmask, t);
    ve[3 * a] = t[0];
    ve[3 * a + 1] = t[1];
    ve[3 * a + 2] = t[2];
  }
  const double *K = K0 + 576 * cls[e];
  double s = 0.0;
  for (int i = 0; i < 24; i++) {
    double t = 0.0;
    for (int j = 0; j < 24; j++) t += K[24 * i + j] * ue[j];
    s += ve[i] * t;
  }
  out[e] = s;
}

void launch_elem_energy2(int nelem, const int *conn, const int *dep_ptr, const int *dep_conn,
                         const float *dep_w, const uint8_t *bcmask, const int *cls,
                         const double *K0, const double *u, const double *v, double *out,
                         hipStream_t st) {
  if (nelem == 0) return;
  hipLaunchKernelGGL(k_elem_energy2, dim3((nelem + 255) / 256), dim3(256), 0, st, nelem, conn,
                     dep_ptr, dep_conn, dep_w, bcmask, cls, K0, u, v, out);
  TMRX_LAUNCH_CHECK();
}

__global__ void k_adjoint_grad(int n, const double *__restrict__ rho, const int16_t *__restrict__ lev,
                               double q, int ptype, const double *__restrict__ prod,
                               const double *__restrict__ ex, double *__restrict__ g) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double r = rho[i];
  double dp;
  if (ptype == 0) {
    dp = (q > 1.0) ? q * pow(r, q - 1.0) : 1.0;
  } else {
    const double t = 1.0 + q * (1.0 - r);
    dp = (q + 1.0) / (t * t);
  }
  g[i] = ex[i] - ldexp(dp, -(int)lev[i]) * prod[i];
}

void launch_adjoint_grad(int n, const double *rho, const int16_t *lev, double q, int ptype,
                         const double *prod, const double *explicit_part, double *g,
                         hipStream_t st) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_adjoint_grad, dim3((n + 255) / 256), dim3(256), 0, st, n, rho, lev, q, ptype,
                     prod, explicit_part, g);
  TMRX_LAUNCH_CHECK();
}

// explicit instantiations
#define TMRX_INST(T)                                                                            \
  template void launch_elem_apply<T>(const ElemBatches &, const T *, const T *, const T *, T *, \
                                     hipStream_t);                                              \
  template void launch_elem_diag<T>(const ElemBatches &, const T *, const T *, T *,             \
                                    hipStream_t, bool);                                         \
  template void launch_assemble<T>(const AsmCSR &, const uint8_t *, const T *, int,             \
                                   const AsmArgs &, hipStream_t);                               \
  template void launch_prolong_add<T>(const TransferCSR &, const uint8_t *, const T *, T *,     \
                                      hipStream_t, int, int64_t, int64_t);                      \
  template void launch_restrict<T>(const TransferCSR &, const uint8_t *, const T *, T *,        \
                                   hipStream_t, int, int64_t, int64_t, const T *, double, T *); \
  template void launch_cheb_first<T>(int, const T *, const T *, double, T *, T *,               \
                                     hipStream_t, int, const int *);                            \
  template void launch_dense_solve<T>(const CoarseBlocks &, const T *, T *, hipStream_t, int);  \
  template void launch_fill<T>(int64_t, T, T *, hipStream_t);                                 \
  template void launch_tile_apply<T>(const Tiles &, const T *, const T *, const T *, int,       \
                                     const AsmArgs &, hipStream_t, int, int64_t);
TMRX_INST(float)
TMRX_INST(double)
template void launch_convert<float, double>(int64_t, const float *, double *, hipStream_t);
template void launch_convert<double, float>(int64_t, const double *, float *, hipStream_t);

}  // namespace tmrx
