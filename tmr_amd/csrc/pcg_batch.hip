// Batched FP64 PCG vector kernels: nrhs right-hand sides (grid y) of one
// finest-level vector each, every vector split into independent components
// ("segments") that keep their own PCG scalars and stop on their own
// convergence.  Reductions are deterministic: fixed blocks, fixed-order sums.
#include <hip/hip_runtime.h>

#include "kernels.h"

namespace tmrx {

namespace {
constexpr int BB = 256;  // threads per vector block

__device__ __forceinline__ double bsum(double v, double *red) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[w] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int i = 0; i < (int)(blockDim.x >> 6); i++) s += red[i];
  return s;
}
}  // namespace

// partials[rhs * nblk + b] = a . b over block b (0 for stopped segments when
// act is given)
__global__ __launch_bounds__(BB) void k_b_dot1(SegPlan pl, const double *__restrict__ a,
                                               const double *__restrict__ b,
                                               const int *__restrict__ act,
                                               double *__restrict__ partials) {
  __shared__ double red[BB / 64];
  const int blk = blockIdx.x, rhs = blockIdx.y;
  const int sidx = rhs * pl.nseg + pl.blk_seg[blk];
  const int64_t off = (int64_t)rhs * pl.n;
  double s = 0.0;
  if (!act || act[sidx])
    for (int64_t i = pl.blk_lo[blk] + threadIdx.x; i < pl.blk_lo[blk + 1]; i += BB)
      s += a[off + i] * b[off + i];
  const double t = bsum(s, red);
  if (threadIdx.x == 0) partials[(int64_t)rhs * pl.nblk + blk] = t;
}

// r -= alpha q (and its FP32 copy), partial |r|^2; x += alpha p is left to
// the p update of the same iteration (it reads p there anyway)
__global__ __launch_bounds__(BB) void k_b_update_xr(SegPlan pl, double *__restrict__ r,
                                                    const double *__restrict__ q,
                                                    float *__restrict__ r32,
                                                    const double *__restrict__ sc,
                                                    const int *__restrict__ act,
                                                    double *__restrict__ partials,
                                                    const float *__restrict__ dinv32, double c0,
                                                    float *__restrict__ x1) {
  __shared__ double red[BB / 64];
  const int blk = blockIdx.x, rhs = blockIdx.y;
  const int sidx = rhs * pl.nseg + pl.blk_seg[blk];
  const int64_t off = (int64_t)rhs * pl.n;
  double s = 0.0;
  if (act[sidx]) {
    const double alpha = sc[(int64_t)sidx * SC_N + SC_ALPHA];
    for (int64_t i = pl.blk_lo[blk] + threadIdx.x; i < pl.blk_lo[blk + 1]; i += BB) {
      const int64_t k = off + i;
      const double rn = r[k] - alpha * q[k];
      r[k] = rn;
      if (r32) {
        const float rf = (float)rn;
        r32[k] = rf;
        if (x1) x1[k] = (float)c0 * dinv32[i] * rf;
      }
      s += rn * rn;
    }
  } else if (r32) {
    for (int64_t i = pl.blk_lo[blk] + threadIdx.x; i < pl.blk_lo[blk + 1]; i += BB) {
      const float rf = (float)r[off + i];
      r32[off + i] = rf;
      if (x1) x1[off + i] = (float)c0 * dinv32[i] * rf;
    }
  }
  const double t = bsum(s, red);
  if (threadIdx.x == 0) partials[(int64_t)rhs * pl.nblk + blk] = t;
}

// partials z.r and z.q (z in FP64, or straight from the FP32 V-cycle)
template <class TZ>
__global__ __launch_bounds__(BB) void k_b_dot2(SegPlan pl, const TZ *__restrict__ z,
                                               const double *__restrict__ r,
                                               const double *__restrict__ q,
                                               const int *__restrict__ act,
                                               double *__restrict__ partials) {
  __shared__ double red[BB / 64];
  const int blk = blockIdx.x, rhs = blockIdx.y;
  const int sidx = rhs * pl.nseg + pl.blk_seg[blk];
  const int64_t off = (int64_t)rhs * pl.n;
  double a = 0.0, c = 0.0;
  if (act[sidx])
    for (int64_t i = pl.blk_lo[blk] + threadIdx.x; i < pl.blk_lo[blk + 1]; i += BB) {
      const double zv = (double)z[off + i];
      a += zv * r[off + i];
      c += zv * q[off + i];
    }
  const double sa = bsum(a, red);
  const double sb = bsum(c, red);
  if (threadIdx.x == 0) {
    partials[2 * ((int64_t)rhs * pl.nblk + blk)] = sa;
    partials[2 * ((int64_t)rhs * pl.nblk + blk) + 1] = sb;
  }
}

// x += alpha p for a segment whose alpha was formed in the iteration before
// *iter (also the iteration at which it stopped), then p = z + beta p while
// it runs
template <class TZ>
__global__ __launch_bounds__(BB) void k_b_update_p(SegPlan pl, double *__restrict__ p,
                                                   const TZ *__restrict__ z,
                                                   const double *__restrict__ sc,
                                                   const int *__restrict__ act,
                                                   double *__restrict__ x,
                                                   const int *__restrict__ iter) {
  const int blk = blockIdx.x, rhs = blockIdx.y;
  const int sidx = rhs * pl.nseg + pl.blk_seg[blk];
  const double *s = sc + (int64_t)sidx * SC_N;
  const bool upx = s[SC_XIT] == (double)(*iter - 1);
  const bool upp = act[sidx] != 0;
  if (!upx && !upp) return;
  const double alpha = s[SC_ALPHA], beta = s[SC_BETA];
  const int64_t off = (int64_t)rhs * pl.n;
  for (int64_t i = pl.blk_lo[blk] + threadIdx.x; i < pl.blk_lo[blk + 1]; i += BB) {
    const double pv = p[off + i];
    if (upx) x[off + i] += alpha * pv;
    if (upp) p[off + i] = (double)z[off + i] + beta * pv;
  }
}

// copy a into b for running segments (b = a everywhere when act is null)
__global__ __launch_bounds__(BB) void k_b_copy(SegPlan pl, const double *__restrict__ a,
                                               double *__restrict__ b,
                                               const int *__restrict__ act) {
  const int blk = blockIdx.x, rhs = blockIdx.y;
  if (act && !act[rhs * pl.nseg + pl.blk_seg[blk]]) return;
  const int64_t off = (int64_t)rhs * pl.n;
  for (int64_t i = pl.blk_lo[blk] + threadIdx.x; i < pl.blk_lo[blk + 1]; i += BB)
    b[off + i] = a[off + i];
}

// one block per (segment, right-hand side): fixed-order sum of the
// segment's partials, then the PCG scalar update of `mode`
__global__ __launch_bounds__(256) void k_b_finalize(SegPlan pl, int mode,
                                                    const double *__restrict__ partials,
                                                    double *__restrict__ sc, int *__restrict__ act,
                                                    double *__restrict__ thr2,
                                                    const int *__restrict__ iter,
                                                    int *__restrict__ itconv, double rtol2,
                                                    double atol2) {
  __shared__ double red[4];
  const int seg = blockIdx.x, rhs = blockIdx.y;
  const int sidx = rhs * pl.nseg + seg;
  const bool two = mode == FIN_BETA;
  int lo, hi;
  const double *part;
  if (mode == FIN_PQ) {
    lo = pl.seg_tile[seg];
    hi = pl.seg_tile[seg + 1];
    part = partials + (int64_t)rhs * pl.ntiles;
  } else {
    lo = pl.seg_blk[seg];
    hi = pl.seg_blk[seg + 1];
    part = partials + (two ? 2 : 1) * (int64_t)rhs * pl.nblk;
  }
  // one block does the whole sum: eight partials per thread are loaded
  // before any is added (same order of additions as one at a time)
  double a = 0.0, b = 0.0;
  constexpr int U = 8;
  for (int i0 = lo + threadIdx.x; i0 < hi; i0 += U * 256) {
    double va[U], vb[U];
#pragma unroll
    for (int k = 0; k < U; k++) {
      const int i = min(i0 + k * 256, hi - 1);
      va[k] = two ? part[2 * i] : part[i];
      vb[k] = two ? part[2 * i + 1] : 0.0;
    }
#pragma unroll
    for (int k = 0; k < U; k++)
      if (i0 + k * 256 < hi) {
        a += va[k];
        b += vb[k];
      }
  }
  const double sa = bsum(a, red);
  const double sb = two ? bsum(b, red) : 0.0;
  if (threadIdx.x != 0) return;
  double *s = sc + (int64_t)sidx * SC_N;
  switch (mode) {
    case FIN_BB:  // |b|^2 of the clamped right-hand side: threshold, start or stop
      s[SC_BB] = sa;
      s[SC_RR] = sa;
      s[SC_XIT] = -1.0;
      thr2[sidx] = fmax(rtol2 * sa, atol2);
      act[sidx] = (sa > 0.0 && sa > atol2) ? 1 : 0;
      itconv[sidx] = 0;
      break;
    case FIN_RZ:
      if (act[sidx]) s[SC_RZ] = sa;
      break;
    case FIN_PQ:
      if (act[sidx]) {
        s[SC_PQ] = sa;
        s[SC_ALPHA] = s[SC_RZ] / sa;
        s[SC_XIT] = (double)*iter;  // x += alpha p is due in this iteration
      }
      break;
    case FIN_RR:
      if (act[sidx]) {
        s[SC_RR] = sa;
        const int it = *iter;
        if (!(sa == sa)) {  // breakdown: stop, flagged by a negative count
          act[sidx] = 0;
          itconv[sidx] = -it;
        } else if (sa <= thr2[sidx]) {
          act[sidx] = 0;
          itconv[sidx] = it;
        } else {
          itconv[sidx] = it;
        }
      }
      break;
    case FIN_BETA:  // flexible (Polak-Ribiere) beta = z_{k+1}.(r_{k+1}-r_k) / z_k.r_k
      if (act[sidx]) {
        s[SC_RZNEW] = sa;
        s[SC_ZQ] = sb;
        s[SC_BETA] = -s[SC_ALPHA] * sb / s[SC_RZ];
        s[SC_RZ] = sa;
      }
      break;
    case FIN_TRUE:
      s[SC_TRR] = sa;
      break;
  }
}

__global__ void k_b_tick(int *__restrict__ iter, const int *__restrict__ act, int nidx,
                         int *__restrict__ nact) {
  // one wave: count running segments, advance the iteration counter
  int c = 0;
  for (int i = threadIdx.x; i < nidx; i += 64) c += act[i] ? 1 : 0;
  for (int off = 32; off > 0; off >>= 1) c += __shfl_down(c, off, 64);
  if (threadIdx.x == 0) {
    *nact = c;
    *iter += 1;
  }
}

void launch_b_dot1(const SegPlan &pl, const double *a, const double *b, const int *act,
                   double *partials, hipStream_t st) {
  hipLaunchKernelGGL(k_b_dot1, dim3(pl.nblk, pl.nrhs), dim3(BB), 0, st, pl, a, b, act, partials);
  TMRX_LAUNCH_CHECK();
}
void launch_b_update_xr(const SegPlan &pl, double *r, const double *q, float *r32,
                        const double *sc, const int *act, double *partials, hipStream_t st,
                        const float *dinv32, double c0, float *x1) {
  hipLaunchKernelGGL(k_b_update_xr, dim3(pl.nblk, pl.nrhs), dim3(BB), 0, st, pl, r, q, r32, sc,
                     act, partials, dinv32, c0, x1);
  TMRX_LAUNCH_CHECK();
}
template <class TZ>
void launch_b_dot2(const SegPlan &pl, const TZ *z, const double *r, const double *q,
                   const int *act, double *partials, hipStream_t st) {
  hipLaunchKernelGGL(k_b_dot2<TZ>, dim3(pl.nblk, pl.nrhs), dim3(BB), 0, st, pl, z, r, q, act,
                     partials);
  TMRX_LAUNCH_CHECK();
}
template <class TZ>
void launch_b_update_p(const SegPlan &pl, double *p, const TZ *z, const double *sc,
                       const int *act, double *x, const int *iter, hipStream_t st) {
  hipLaunchKernelGGL(k_b_update_p<TZ>, dim3(pl.nblk, pl.nrhs), dim3(BB), 0, st, pl, p, z, sc,
                     act, x, iter);
  TMRX_LAUNCH_CHECK();
}
template void launch_b_dot2<float>(const SegPlan &, const float *, const double *, const double *,
                                   const int *, double *, hipStream_t);
template void launch_b_dot2<double>(const SegPlan &, const double *, const double *,
                                    const double *, const int *, double *, hipStream_t);
template void launch_b_update_p<float>(const SegPlan &, double *, const float *, const double *,
                                       const int *, double *, const int *, hipStream_t);
template void launch_b_update_p<double>(const SegPlan &, double *, const double *,
                                        const double *, const int *, double *, const int *,
                                        hipStream_t);
void launch_b_copy(const SegPlan &pl, const double *a, double *b, const int *act,
                   hipStream_t st) {
  hipLaunchKernelGGL(k_b_copy, dim3(pl.nblk, pl.nrhs), dim3(BB), 0, st, pl, a, b, act);
  TMRX_LAUNCH_CHECK();
}
void launch_b_finalize(const SegPlan &pl, int mode, const double *partials, double *sc, int *act,
                       double *thr2, const int *iter, int *itconv, double rtol2, double atol2,
                       hipStream_t st) {
  hipLaunchKernelGGL(k_b_finalize, dim3(pl.nseg, pl.nrhs), dim3(256), 0, st, pl, mode, partials,
                     sc, act, thr2, iter, itconv, rtol2, atol2);
  TMRX_LAUNCH_CHECK();
}
void launch_b_tick(int *iter, const int *act, int nidx, int *nact, hipStream_t st) {
  hipLaunchKernelGGL(k_b_tick, dim3(1), dim3(64), 0, st, iter, act, nidx, nact);
  TMRX_LAUNCH_CHECK();
}

}  // namespace tmrx
