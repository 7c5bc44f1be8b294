/*
  Oracle multigrid (test infrastructure and CPU baseline only; nothing in the
  product loads this file).

  A CPU restatement, on assembled FP64 CSR matrices, of the two solvers the
  reference builds for TMRTopoProblem:

  * TACSMg as TMR_CreateTACSMg configures it (src/TMR_RefinementTools.cpp:37-107):
    one interpolation per level (createInterpolation), Galerkin coarse
    operators A_{l+1} = P^T A_l P (use_galerkin), symmetric SOR smoothing
    with omega = 1 and one sweep, or Chebyshev smoothing on D^-1 A, and a
    direct solve of the coarsest level or (use_coarse_direct_solve = 0,
    :87-102, orc_mg_set_coarse) one application of the level's smoother
    there;
  * the Krylov method around it: GMRES with a subspace of 50 and restarts
    (TMRTopoProblem, src/topology/TMRTopoProblem.cpp:197-204), or the
    flexible PCG of the GPU path.

  Clamped dofs: P is masked to free dofs on both sides (P~ = M_f P M_c) and
  the coarse operators get identity rows and columns there.  OpenMP threads
  (orc_set_threads) share the sparse products and vector updates; SOR
  sweeps are sequential.
*/
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

void orc_set_threads(int n);
int orc_get_threads(void);

typedef struct {
  int n, m; /* rows, columns */
  int *rowp, *cols;
  double *vals;
} mat_t;

#define MG_MAX 16
static struct {
  int nl;
  mat_t A[MG_MAX], P[MG_MAX], R[MG_MAX];
  uint8_t *free_[MG_MAX]; /* 1 on free dofs */
  double *dinv[MG_MAX];
  double lam[MG_MAX];
  double *Lc; /* dense Cholesky factor of the coarsest level (lower, row-major) */
  double *w1[MG_MAX], *w2[MG_MAX], *w3[MG_MAX], *w4[MG_MAX];
  int *order[MG_MAX]; /* SOR node order per level (NULL: natural) */
  int smoother; /* 0 Chebyshev, 1 symmetric SOR */
  int coarse_smooth; /* 1: the coarsest level is smoothed, not solved */
  int degree;
  double lower, upper, omega;
} g;

static void mat_free(mat_t *a) {
  free(a->rowp);
  free(a->cols);
  free(a->vals);
  memset(a, 0, sizeof(*a));
}

void orc_mg_free(void) {
  for (int l = 0; l < MG_MAX; l++) {
    mat_free(&g.A[l]);
    mat_free(&g.P[l]);
    mat_free(&g.R[l]);
    free(g.free_[l]);
    free(g.dinv[l]);
    free(g.w1[l]);
    free(g.w2[l]);
    free(g.w3[l]);
    free(g.w4[l]);
    free(g.order[l]);
    g.order[l] = NULL;
    g.free_[l] = NULL;
    g.dinv[l] = g.w1[l] = g.w2[l] = g.w3[l] = g.w4[l] = NULL;
  }
  free(g.Lc);
  g.Lc = NULL;
  g.nl = 0;
}

/* level l's operator (the finest from orc_assemble's CSR) and its clamped
   dofs (bit d of bcmask[node] = dof d clamped) */
int orc_mg_set_level(int l, int nnodes, const uint8_t *bcmask) {
  if (l < 0 || l >= MG_MAX) return -1;
  if (l == 0) orc_mg_free();
  if (l + 1 > g.nl) g.nl = l + 1;
  const int n = 3 * nnodes;
  free(g.free_[l]);
  g.free_[l] = (uint8_t *)malloc(n);
  for (int i = 0; i < nnodes; i++)
    for (int d = 0; d < 3; d++) g.free_[l][3 * i + d] = !((bcmask[i] >> d) & 1);
  g.A[l].n = g.A[l].m = n;
  return 0;
}

int orc_mg_set_matrix(int l, const int *rowp, const int *cols, const double *vals) {
  mat_t *a = &g.A[l];
  const int n = a->n;
  free(a->rowp);
  free(a->cols);
  free(a->vals);
  a->rowp = (int *)malloc(sizeof(int) * (n + 1));
  memcpy(a->rowp, rowp, sizeof(int) * (n + 1));
  const int nnz = rowp[n];
  a->cols = (int *)malloc(sizeof(int) * (nnz ? nnz : 1));
  a->vals = (double *)malloc(sizeof(double) * (nnz ? nnz : 1));
  memcpy(a->cols, cols, sizeof(int) * nnz);
  memcpy(a->vals, vals, sizeof(double) * nnz);
  return 0;
}

/* prolongation from level l+1 to level l (node-level CSR from
   createInterpolation: rows fine nodes, columns coarse nodes), expanded to
   the 3 dofs and masked: P~ = M_f (P x I3) M_c; R = P~^T */
int orc_mg_set_interp(int l, int nf, int nc, const int *rowp, const int *cols, const double *w) {
  mat_t *P = &g.P[l], *R = &g.R[l];
  mat_free(P);
  mat_free(R);
  const uint8_t *ff = g.free_[l], *fc = g.free_[l + 1];
  P->n = 3 * nf;
  P->m = 3 * nc;
  P->rowp = (int *)calloc(P->n + 1, sizeof(int));
  const int nnz0 = rowp[nf];
  P->cols = (int *)malloc(sizeof(int) * 3 * (nnz0 ? nnz0 : 1));
  P->vals = (double *)malloc(sizeof(double) * 3 * (nnz0 ? nnz0 : 1));
  int k = 0;
  for (int i = 0; i < nf; i++)
    for (int d = 0; d < 3; d++) {
      const int r = 3 * i + d;
      if (ff[r])
        for (int j = rowp[i]; j < rowp[i + 1]; j++) {
          const int c = 3 * cols[j] + d;
          if (!fc[c]) continue;
          P->cols[k] = c;
          P->vals[k++] = w[j];
        }
      P->rowp[r + 1] = k;
    }
  /* transpose */
  R->n = P->m;
  R->m = P->n;
  R->rowp = (int *)calloc(R->n + 1, sizeof(int));
  R->cols = (int *)malloc(sizeof(int) * (k ? k : 1));
  R->vals = (double *)malloc(sizeof(double) * (k ? k : 1));
  for (int j = 0; j < k; j++) R->rowp[P->cols[j] + 1]++;
  for (int i = 0; i < R->n; i++) R->rowp[i + 1] += R->rowp[i];
  int *fill = (int *)malloc(sizeof(int) * (R->n + 1));
  memcpy(fill, R->rowp, sizeof(int) * (R->n + 1));
  for (int r = 0; r < P->n; r++)
    for (int j = P->rowp[r]; j < P->rowp[r + 1]; j++) {
      const int c = P->cols[j];
      R->cols[fill[c]] = r;
      R->vals[fill[c]++] = P->vals[j];
    }
  free(fill);
  return 0;
}

/* Z = X Y (Gustavson, columns of each row in ascending order) */
static void spgemm(const mat_t *X, const mat_t *Y, mat_t *Z) {
  const int n = X->n, m = Y->m;
  int *mark = (int *)malloc(sizeof(int) * m);
  double *acc = (double *)calloc(m, sizeof(double));
  for (int j = 0; j < m; j++) mark[j] = -1;
  int cap = 16 * (n + 1), nnz = 0;
  int *rowp = (int *)malloc(sizeof(int) * (n + 1));
  int *cols = (int *)malloc(sizeof(int) * cap);
  double *vals = (double *)malloc(sizeof(double) * cap);
  int *lst = (int *)malloc(sizeof(int) * m);
  rowp[0] = 0;
  for (int i = 0; i < n; i++) {
    int cnt = 0;
    for (int a = X->rowp[i]; a < X->rowp[i + 1]; a++) {
      const int k = X->cols[a];
      const double xv = X->vals[a];
      for (int b = Y->rowp[k]; b < Y->rowp[k + 1]; b++) {
        const int j = Y->cols[b];
        if (mark[j] != i) {
          mark[j] = i;
          lst[cnt++] = j;
          acc[j] = 0.0;
        }
        acc[j] += xv * Y->vals[b];
      }
    }
    /* insertion sort of the row's columns (rows are short) */
    for (int p = 1; p < cnt; p++) {
      const int v = lst[p];
      int q = p - 1;
      while (q >= 0 && lst[q] > v) {
        lst[q + 1] = lst[q];
        q--;
      }
      lst[q + 1] = v;
    }
    if (nnz + cnt > cap) {
      while (nnz + cnt > cap) cap *= 2;
      cols = (int *)realloc(cols, sizeof(int) * cap);
      vals = (double *)realloc(vals, sizeof(double) * cap);
    }
    for (int p = 0; p < cnt; p++) {
      cols[nnz] = lst[p];
      vals[nnz++] = acc[lst[p]];
    }
    rowp[i + 1] = nnz;
  }
  free(mark);
  free(acc);
  free(lst);
  Z->n = n;
  Z->m = m;
  Z->rowp = rowp;
  Z->cols = cols;
  Z->vals = vals;
}

/* A_{l+1} = R A_l P + I on the clamped dofs of level l+1 */
int orc_mg_galerkin(int l) {
  mat_t AP, C;
  spgemm(&g.A[l], &g.P[l], &AP);
  spgemm(&g.R[l], &AP, &C);
  mat_free(&AP);
  const uint8_t *fc = g.free_[l + 1];
  for (int i = 0; i < C.n; i++)
    if (!fc[i]) {
      /* the masked product leaves the row and column of a clamped dof empty */
      int has = 0;
      for (int k = C.rowp[i]; k < C.rowp[i + 1]; k++)
        if (C.cols[k] == i) {
          C.vals[k] = 1.0;
          has = 1;
        }
      if (!has) {
        /* insert the diagonal entry */
        const int nnz = C.rowp[C.n];
        C.cols = (int *)realloc(C.cols, sizeof(int) * (nnz + 1));
        C.vals = (double *)realloc(C.vals, sizeof(double) * (nnz + 1));
        int pos = C.rowp[i];
        while (pos < C.rowp[i + 1] && C.cols[pos] < i) pos++;
        memmove(&C.cols[pos + 1], &C.cols[pos], sizeof(int) * (nnz - pos));
        memmove(&C.vals[pos + 1], &C.vals[pos], sizeof(double) * (nnz - pos));
        C.cols[pos] = i;
        C.vals[pos] = 1.0;
        for (int r = i + 1; r <= C.n; r++) C.rowp[r]++;
      }
    }
  mat_free(&g.A[l + 1]);
  g.A[l + 1] = C;
  return C.rowp[C.n];
}

/* copy out level l's operator: sizes when the arrays are NULL */
int orc_mg_get_matrix(int l, int *n, int *nnz, int *rowp, int *cols, double *vals) {
  const mat_t *a = &g.A[l];
  if (n) *n = a->n;
  if (nnz) *nnz = a->rowp ? a->rowp[a->n] : 0;
  if (rowp) memcpy(rowp, a->rowp, sizeof(int) * (a->n + 1));
  if (cols) memcpy(cols, a->cols, sizeof(int) * a->rowp[a->n]);
  if (vals) memcpy(vals, a->vals, sizeof(double) * a->rowp[a->n]);
  return 0;
}

static void spmv(const mat_t *A, const double *x, double *y) {
  const int nt = orc_get_threads();
#pragma omp parallel for schedule(static) num_threads(nt)
  for (int i = 0; i < A->n; i++) {
    double s = 0.0;
    for (int k = A->rowp[i]; k < A->rowp[i + 1]; k++) s += A->vals[k] * x[A->cols[k]];
    y[i] = s;
  }
}

static double dot(int n, const double *a, const double *b) {
  double s = 0.0;
  const int nt = orc_get_threads();
#pragma omp parallel for schedule(static) reduction(+ : s) num_threads(nt)
  for (int i = 0; i < n; i++) s += a[i] * b[i];
  return s;
}

/* the coarsest level solved directly (direct = 1, the default) or smoothed
   like the others (TMR_CreateTACSMg, src/TMR_RefinementTools.cpp:83-102);
   takes effect at the next orc_mg_setup */
void orc_mg_set_coarse(int direct) { g.coarse_smooth = !direct; }

/* smoother set-up: Jacobi diagonal and lambda_max(D^-1 A) by 10 power
   iterations on every smoothed level; dense Cholesky of the coarsest level
   when it is solved directly */
int orc_mg_setup(int smoother, int degree, double lower, double upper, double omega) {
  g.smoother = smoother;
  g.degree = degree < 1 ? 1 : degree;
  g.lower = lower;
  g.upper = upper;
  g.omega = omega;
  for (int l = 0; l < g.nl; l++) {
    const int n = g.A[l].n;
    free(g.dinv[l]);
    free(g.w1[l]);
    free(g.w2[l]);
    free(g.w3[l]);
    free(g.w4[l]);
    g.dinv[l] = (double *)malloc(sizeof(double) * n);
    g.w1[l] = (double *)calloc(n, sizeof(double));
    g.w2[l] = (double *)calloc(n, sizeof(double));
    g.w3[l] = (double *)calloc(n, sizeof(double));
    g.w4[l] = (double *)calloc(n, sizeof(double));
    const mat_t *A = &g.A[l];
    for (int i = 0; i < n; i++) {
      double d = 0.0;
      for (int k = A->rowp[i]; k < A->rowp[i + 1]; k++)
        if (A->cols[k] == i) d = A->vals[k];
      if (!(d > 0.0)) return -1 - l;
      g.dinv[l][i] = 1.0 / d;
    }
    if (l + 1 < g.nl || g.coarse_smooth) {
      double *x = g.w1[l], *y = g.w2[l];
      uint64_t s = 0x2545f4914f6cdd1dULL ^ (uint64_t)(l + 1);
      for (int i = 0; i < n; i++) {
        s ^= s << 13;
        s ^= s >> 7;
        s ^= s << 17;
        x[i] = g.free_[l][i] ? 0.5 + (double)(s >> 11) * 0x1.0p-53 : 0.0;
      }
      double lam = 0.0;
      for (int it = 0; it < 10; it++) {
        spmv(A, x, y);
        for (int i = 0; i < n; i++) y[i] *= g.dinv[l][i];
        lam = sqrt(dot(n, y, y) / dot(n, x, x));
        double *t = x;
        x = y;
        y = t;
      }
      g.lam[l] = lam;
    }
  }
  /* coarsest level: dense Cholesky */
  const int lc = g.nl - 1, m = g.A[lc].n;
  free(g.Lc);
  g.Lc = NULL;
  if (g.coarse_smooth) return 0;
  g.Lc = (double *)calloc((size_t)m * m, sizeof(double));
  for (int i = 0; i < m; i++)
    for (int k = g.A[lc].rowp[i]; k < g.A[lc].rowp[i + 1]; k++)
      g.Lc[(size_t)i * m + g.A[lc].cols[k]] = g.A[lc].vals[k];
  for (int j = 0; j < m; j++) {
    double *Lj = &g.Lc[(size_t)j * m];
    double d = Lj[j];
    for (int k = 0; k < j; k++) d -= Lj[k] * Lj[k];
    if (!(d > 0.0)) return -100;
    d = sqrt(d);
    Lj[j] = d;
    for (int i = j + 1; i < m; i++) {
      double *Li = &g.Lc[(size_t)i * m];
      double t = Li[j];
      for (int k = 0; k < j; k++) t -= Li[k] * Lj[k];
      Li[j] = t / d;
    }
  }
  return 0;
}

static void coarse_solve(const double *b, double *x) {
  const int m = g.A[g.nl - 1].n;
  const double *L = g.Lc;
  for (int i = 0; i < m; i++) {
    double t = b[i];
    for (int k = 0; k < i; k++) t -= L[(size_t)i * m + k] * x[k];
    x[i] = t / L[(size_t)i * m + i];
  }
  for (int i = m - 1; i >= 0; i--) {
    double t = x[i];
    for (int k = i + 1; k < m; k++) t -= L[(size_t)k * m + i] * x[k];
    x[i] = t / L[(size_t)i * m + i];
  }
}

/* the order of level l's SOR sweeps: the rows of node order[0], then of
   order[1], ... (each node's 3 dofs in turn); NULL restores row order */
int orc_mg_set_order(int l, int nnodes, const int *order) {
  if (l < 0 || l >= MG_MAX) return -1;
  free(g.order[l]);
  g.order[l] = NULL;
  if (!order) return 0;
  if (3 * nnodes != g.A[l].n) return -2;
  g.order[l] = (int *)malloc(sizeof(int) * (size_t)nnodes);
  memcpy(g.order[l], order, sizeof(int) * (size_t)nnodes);
  return 0;
}

/* one forward and one backward SOR sweep on A x = b (x updated in place),
   rows in the level's order (orc_mg_set_order); the backward sweep visits
   them in reverse */
static void sor_symmetric(int l, const double *b, double *x) {
  const mat_t *A = &g.A[l];
  const double w = g.omega;
  const int *ord = g.order[l];
  for (int pass = 0; pass < 2; pass++)
    for (int t = 0; t < A->n; t++) {
      const int u = pass == 0 ? t : A->n - 1 - t;
      const int i = ord ? 3 * ord[u / 3] + u % 3 : u;
      double s = b[i], d = 1.0;
      for (int k = A->rowp[i]; k < A->rowp[i + 1]; k++) {
        const int j = A->cols[k];
        if (j == i)
          d = A->vals[k];
        else
          s -= A->vals[k] * x[j];
      }
      x[i] = (1.0 - w) * x[i] + w * s / d;
    }
}

/* Chebyshev steps on D^-1 A (the GPU smoother's recurrence): from x (zero
   when first is set) with k steps */
static void chebyshev(int l, const double *b, double *x, int k, int first) {
  const int n = g.A[l].n;
  const double lmax = g.upper * g.lam[l], lmin = g.lower * g.lam[l];
  const double theta = 0.5 * (lmax + lmin), delta = 0.5 * (lmax - lmin);
  const double sigma = theta / delta;
  double *r = g.w3[l], *d = g.w4[l];
  double rho_prev = 1.0 / sigma;
  const int nt = orc_get_threads();
  for (int j = 0; j < k; j++) {
    if (j == 0 && first) {
      for (int i = 0; i < n; i++) r[i] = b[i];
    } else {
      spmv(&g.A[l], x, r);
#pragma omp parallel for schedule(static) num_threads(nt)
      for (int i = 0; i < n; i++) r[i] = b[i] - r[i];
    }
    if (j == 0) {
      const double c = 1.0 / theta;
#pragma omp parallel for schedule(static) num_threads(nt)
      for (int i = 0; i < n; i++) {
        d[i] = c * g.dinv[l][i] * r[i];
        x[i] = (first ? 0.0 : x[i]) + d[i];
      }
    } else {
      const double rho = 1.0 / (2.0 * sigma - rho_prev);
      const double ca = rho * rho_prev, cc = 2.0 * rho / delta;
#pragma omp parallel for schedule(static) num_threads(nt)
      for (int i = 0; i < n; i++) {
        d[i] = ca * d[i] + cc * g.dinv[l][i] * r[i];
        x[i] += d[i];
      }
      rho_prev = rho;
    }
  }
}

/* the coarsest level's solve: the direct solve, or its smoother once from
   zero */
static void coarse_apply(const double *b, double *x) {
  const int l = g.nl - 1;
  if (!g.coarse_smooth) {
    coarse_solve(b, x);
  } else if (g.smoother == 0) {
    chebyshev(l, b, x, g.degree, 1);
  } else {
    memset(x, 0, sizeof(double) * g.A[l].n);
    sor_symmetric(l, b, x);
  }
}

/* x = V-cycle(b) from zero at level l */
static void vcycle(int l, const double *b, double *x) {
  const int n = g.A[l].n;
  if (l == g.nl - 1) {
    coarse_apply(b, x);
    return;
  }
  if (g.smoother == 0) {
    chebyshev(l, b, x, g.degree, 1);
  } else {
    memset(x, 0, sizeof(double) * n);
    sor_symmetric(l, b, x);
  }
  double *r = g.w1[l];
  spmv(&g.A[l], x, r);
  for (int i = 0; i < n; i++) r[i] = b[i] - r[i];
  const int nc = g.A[l + 1].n;
  double *bc = (double *)malloc(sizeof(double) * nc), *xc = (double *)malloc(sizeof(double) * nc);
  spmv(&g.R[l], r, bc);
  vcycle(l + 1, bc, xc);
  spmv(&g.P[l], xc, r);
  for (int i = 0; i < n; i++) x[i] += r[i];
  free(bc);
  free(xc);
  if (g.smoother == 0)
    chebyshev(l, b, x, g.degree, 0);
  else
    sor_symmetric(l, b, x);
}

/* level l's estimate of lambda_max(D^-1 A); set replaces it (to run the
   smoother with another implementation's estimate) */
double orc_mg_get_lambda(int l) { return (l >= 0 && l < g.nl) ? g.lam[l] : 0.0; }
int orc_mg_set_lambda(int l, double lam) {
  if (l < 0 || l >= g.nl) return -1;
  g.lam[l] = lam;
  return 0;
}

/* one component alone: 0 y = A_l x; 1 y = R_l x; 2 y = P_l x (x on level
   l + 1); 3 y = level l's pre-smoothing from zero with right-hand side x;
   4 y = the coarsest level's solve (direct or smoothed); 5 y = the V-cycle
   from level l */
int orc_mg_op(int op, int l, const double *x, double *y) {
  if (l < 0 || l >= g.nl) return -1;
  const int last = g.nl - 1;
  if ((op >= 1 && op <= 3 && l == last) || (op == 4 && l != last)) return -1;
  switch (op) {
    case 0: spmv(&g.A[l], x, y); break;
    case 1: spmv(&g.R[l], x, y); break;
    case 2: spmv(&g.P[l], x, y); break;
    case 3:
      if (g.smoother == 0) {
        chebyshev(l, x, y, g.degree, 1);
      } else {
        memset(y, 0, sizeof(double) * g.A[l].n);
        sor_symmetric(l, x, y);
      }
      break;
    case 4: coarse_apply(x, y); break;
    case 5: vcycle(l, x, y); break;
    default: return -1;
  }
  return 0;
}

/* flexible PCG (Polak-Ribiere beta) with one V-cycle per iteration, from
   x = 0; b zero on clamped dofs.  Returns iterations (negative if not
   converged) and the recursive relative residual. */
int orc_mg_pcg(const double *b, double *x, double rtol, int maxit, double *relres) {
  const int n = g.A[0].n;
  double *r = (double *)malloc(sizeof(double) * n), *z = (double *)malloc(sizeof(double) * n);
  double *p = (double *)malloc(sizeof(double) * n), *q = (double *)malloc(sizeof(double) * n);
  const int nt = orc_get_threads();
  for (int i = 0; i < n; i++) {
    x[i] = 0.0;
    r[i] = b[i];
  }
  const double bn = sqrt(dot(n, b, b));
  vcycle(0, r, z);
  memcpy(p, z, sizeof(double) * n);
  double rz = dot(n, r, z), rn = bn;
  int it = 0, conv = bn == 0.0;
  while (!conv && it < maxit) {
    it++;
    spmv(&g.A[0], p, q);
    const double alpha = rz / dot(n, p, q);
    double rr = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : rr) num_threads(nt)
    for (int i = 0; i < n; i++) {
      x[i] += alpha * p[i];
      r[i] -= alpha * q[i];
      rr += r[i] * r[i];
    }
    rn = sqrt(rr);
    if (rn <= rtol * bn) {
      conv = 1;
      break;
    }
    vcycle(0, r, z);
    const double rzn = dot(n, r, z), zq = dot(n, z, q);
    const double beta = -alpha * zq / rz;
    rz = rzn;
#pragma omp parallel for schedule(static) num_threads(nt)
    for (int i = 0; i < n; i++) p[i] = z[i] + beta * p[i];
  }
  if (relres) *relres = bn > 0 ? rn / bn : 0.0;
  free(r);
  free(z);
  free(p);
  free(q);
  return conv ? it : -it;
}

/* GMRES(m) with the V-cycle as a right preconditioner, restarted until
   |r| <= rtol |b| or maxit inner iterations (TMRTopoProblem's
   GMRES(50), src/topology/TMRTopoProblem.cpp:197-204). */
int orc_mg_gmres(const double *b, double *x, int m, double rtol, int maxit, double *relres) {
  const int n = g.A[0].n;
  double *V = (double *)malloc(sizeof(double) * (size_t)n * (m + 1));
  double *Zv = (double *)malloc(sizeof(double) * (size_t)n * m);
  double *H = (double *)calloc((size_t)(m + 1) * m, sizeof(double));
  double *cs = (double *)malloc(sizeof(double) * m), *sn = (double *)malloc(sizeof(double) * m);
  double *s = (double *)malloc(sizeof(double) * (m + 1)), *y = (double *)malloc(sizeof(double) * m);
  double *w = (double *)malloc(sizeof(double) * n);
  const int nt = orc_get_threads();
  memset(x, 0, sizeof(double) * n);
  const double bn = sqrt(dot(n, b, b));
  int total = 0, conv = bn == 0.0;
  double rn = bn;
  while (!conv && total < maxit) {
    spmv(&g.A[0], x, w);
    for (int i = 0; i < n; i++) V[i] = b[i] - w[i];
    double beta = sqrt(dot(n, V, V));
    rn = beta;
    if (beta <= rtol * bn) {
      conv = 1;
      break;
    }
    for (int i = 0; i < n; i++) V[i] /= beta;
    memset(s, 0, sizeof(double) * (m + 1));
    s[0] = beta;
    int k = 0;
    for (; k < m && total < maxit; k++, total++) {
      double *zk = &Zv[(size_t)k * n], *vk1 = &V[(size_t)(k + 1) * n];
      vcycle(0, &V[(size_t)k * n], zk);
      spmv(&g.A[0], zk, vk1);
      for (int j = 0; j <= k; j++) {
        const double *vj = &V[(size_t)j * n];
        const double h = dot(n, vk1, vj);
        H[(size_t)j * m + k] = h;
#pragma omp parallel for schedule(static) num_threads(nt)
        for (int i = 0; i < n; i++) vk1[i] -= h * vj[i];
      }
      const double hn = sqrt(dot(n, vk1, vk1));
      H[(size_t)(k + 1) * m + k] = hn;
      if (hn > 0.0)
        for (int i = 0; i < n; i++) vk1[i] /= hn;
      for (int j = 0; j < k; j++) {
        const double a = H[(size_t)j * m + k], c = H[(size_t)(j + 1) * m + k];
        H[(size_t)j * m + k] = cs[j] * a + sn[j] * c;
        H[(size_t)(j + 1) * m + k] = -sn[j] * a + cs[j] * c;
      }
      const double a = H[(size_t)k * m + k], c = H[(size_t)(k + 1) * m + k];
      const double r = sqrt(a * a + c * c);
      cs[k] = r > 0 ? a / r : 1.0;
      sn[k] = r > 0 ? c / r : 0.0;
      H[(size_t)k * m + k] = r;
      H[(size_t)(k + 1) * m + k] = 0.0;
      s[k + 1] = -sn[k] * s[k];
      s[k] = cs[k] * s[k];
      rn = fabs(s[k + 1]);
      if (rn <= rtol * bn) {
        conv = 1;
        k++;
        total++;
        break;
      }
    }
    for (int i = k - 1; i >= 0; i--) {
      double t = s[i];
      for (int j = i + 1; j < k; j++) t -= H[(size_t)i * m + j] * y[j];
      y[i] = t / H[(size_t)i * m + i];
    }
    for (int j = 0; j < k; j++) {
      const double *zj = &Zv[(size_t)j * n];
#pragma omp parallel for schedule(static) num_threads(nt)
      for (int i = 0; i < n; i++) x[i] += y[j] * zj[i];
    }
  }
  if (relres) *relres = bn > 0 ? rn / bn : 0.0;
  free(V);
  free(Zv);
  free(H);
  free(cs);
  free(sn);
  free(s);
  free(y);
  free(w);
  return conv ? total : -total;
}
