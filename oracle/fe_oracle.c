/*
  FP64 CPU oracle for the hex8 SIMP solve path -- TEST INFRASTRUCTURE ONLY.

  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
  load this library, and only as the checker or the timed CPU baseline; the
  product path never calls it.

  It restates, with plain loops and an explicitly assembled sparse matrix,
  what the reference computes through TACS for this path:
    - element stiffness of a trilinear hex, 2x2x2 Gauss, isotropic material
      with engineering shear strains [exx,eyy,ezz,gyz,gxz,gxy]
      (TACSLinearElasticity3D as used by src/topology/TMROctConstitutive.cpp:476-551);
    - SIMP penalty rho^q + k0 and RAMP rho/(1+q(1-rho)) + k0
      (src/topology/TMROctConstitutive.cpp:541-546) and their derivatives
      (src/topology/TMROctConstitutive.cpp:597-619);
    - hanging-node constraints u_dep = sum w u_parent (TACSAssembler
      setDependentNodes, src/TMR_TACSCreator.cpp:405): A = Z^T K Z;
    - clamped dofs: identity rows/columns and zero right-hand side;
    - the solve (TACS GMRES + multigrid in TMRTopoProblem) replaced by Jacobi
      PCG to a tight tolerance: the converged displacement is unique, so any
      convergent method pins it;
    - compliance f^T u (src/topology/TMRTopoProblem.cpp:1034) and its
      derivative -psi^T dK/drho u with psi = u
      (src/topology/TMRTopoProblem.cpp:1275-1280).
  Pins: tests/test_oracle_fe.py (closed-form unit-cube stiffness, rigid-body
  null space, patch test, dense Cholesky agreement, energy identity,
  central differences).
*/
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

static void iso_C(double E, double nu, double C[6][6]) {
  double lam = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu));
  double mu = 0.5 * E / (1.0 + nu);
  memset(C, 0, sizeof(double) * 36);
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) C[i][j] = lam;
  for (int i = 0; i < 3; i++) C[i][i] += 2.0 * mu;
  for (int i = 3; i < 6; i++) C[i][i] = mu;
}

/* gradient of shape function a at (xi,eta,zeta) in physical coordinates */
static void grad_N(int a, const double h[3], const double p[3], double g[3]) {
  int bits[3] = {a & 1, (a >> 1) & 1, (a >> 2) & 1};
  double f[3], df[3];
  for (int d = 0; d < 3; d++) {
    double s = bits[d] ? 1.0 : -1.0;
    f[d] = 0.5 + 0.5 * s * p[d];
    df[d] = s / h[d];
  }
  g[0] = df[0] * f[1] * f[2];
  g[1] = f[0] * df[1] * f[2];
  g[2] = f[0] * f[1] * df[2];
}

/* strain (Voigt, engineering shear) of displacement u_e at point p */
static void strain_at(const double h[3], const double p[3], const double *ue, double e[6]) {
  double G[3][3] = {{0}}; /* G[i][j] = du_i/dx_j */
  for (int a = 0; a < 8; a++) {
    double g[3];
    grad_N(a, h, p, g);
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) G[i][j] += ue[3 * a + i] * g[j];
  }
  e[0] = G[0][0];
  e[1] = G[1][1];
  e[2] = G[2][2];
  e[3] = G[1][2] + G[2][1];
  e[4] = G[0][2] + G[2][0];
  e[5] = G[0][1] + G[1][0];
}

/* Contribution of Gauss point q (bit d of q set: coordinate d at +1/sqrt(3))
   to K, column by column: K_q[:,c] = w detJ B_q^T C (B_q e_c), using the
   strain of unit displacement vectors rather than a stored B matrix */
void orc_hex8_stiffness_gp(double hx, double hy, double hz, double E, double nu, int q,
                           double *K) {
  double C[6][6];
  iso_C(E, nu, C);
  const double h[3] = {hx, hy, hz};
  const double gp = 1.0 / sqrt(3.0);
  const double detJ = hx * hy * hz / 8.0;
  double unit[24];
  double p[3] = {(q & 1) ? gp : -gp, (q & 2) ? gp : -gp, (q & 4) ? gp : -gp};
  double eps[24][6];
  for (int c = 0; c < 24; c++) {
    memset(unit, 0, sizeof(unit));
    unit[c] = 1.0;
    strain_at(h, p, unit, eps[c]);
  }
  for (int r = 0; r < 24; r++)
    for (int c = 0; c < 24; c++) {
      double s = 0.0;
      for (int i = 0; i < 6; i++) {
        double sig = 0.0;
        for (int j = 0; j < 6; j++) sig += C[i][j] * eps[c][j];
        s += eps[r][i] * sig;
      }
      K[24 * r + c] = detJ * s;
    }
}

/* K = sum over the 8 Gauss points */
void orc_hex8_stiffness(double hx, double hy, double hz, double E, double nu, double *K) {
  double Kq[576];
  memset(K, 0, sizeof(double) * 576);
  for (int q = 0; q < 8; q++) {
    orc_hex8_stiffness_gp(hx, hy, hz, E, nu, q, Kq);
    for (int i = 0; i < 576; i++) K[i] += Kq[i];
  }
}

/* element stiffness of a nodal design: sum_q s[q] K_q */
static void hex8_stiffness_scaled(const double h[3], double E, double nu, const double *sq,
                                  double *K) {
  double Kq[576];
  memset(K, 0, sizeof(double) * 576);
  for (int q = 0; q < 8; q++) {
    orc_hex8_stiffness_gp(h[0], h[1], h[2], E, nu, q, Kq);
    for (int i = 0; i < 576; i++) K[i] += sq[q] * Kq[i];
  }
}

double orc_penalty(double rho, double q, double k0, int ptype) {
  if (ptype == 0) return pow(rho, q) + k0;
  return rho / (1.0 + q * (1.0 - rho)) + k0;
}

double orc_dpenalty(double rho, double q, int ptype) {
  if (ptype == 0) return (q > 1.0) ? q * pow(rho, q - 1.0) : 1.0;
  double t = 1.0 + q * (1.0 - rho);
  return (q + 1.0) / (t * t);
}

/* ------------------------------------------------------------------------- */
typedef struct {
  int n;       /* dofs */
  int *rowp;   /* n + 1 */
  int *cols;
  double *vals;
} csr_t;

static csr_t *g_A = NULL;

typedef struct {
  int r, c;
  double v;
} trip_t;

static int trip_cmp(const void *a, const void *b) {
  const trip_t *x = (const trip_t *)a, *y = (const trip_t *)b;
  if (x->r != y->r) return x->r < y->r ? -1 : 1;
  if (x->c != y->c) return x->c < y->c ? -1 : 1;
  return 0;
}

void orc_free(void) {
  if (g_A) {
    free(g_A->rowp);
    free(g_A->cols);
    free(g_A->vals);
    free(g_A);
    g_A = NULL;
  }
}

/*
  Assemble A = Z^T (sum_e s_e K_e) Z over the independent nodes with
  identity rows/columns on clamped dofs.
    nelem, conn[8*nelem] (>= 0 node, < 0 -(dep+1)), dep_ptr/dep_conn/dep_w,
    hdim[3*nelem] element edge lengths, s[nelem] stiffness factor,
    bcmask[nnodes] (bit d = dof d clamped), E, nu.
  Returns the number of stored entries.
*/
static long assemble_impl(int nnodes, int nelem, const int *conn, const int *dep_ptr,
                          const int *dep_conn, const double *dep_w, const double *hdim,
                          const double *s, int per_gp, const uint8_t *bcmask, double E, double nu);

long orc_assemble(int nnodes, int nelem, const int *conn, const int *dep_ptr, const int *dep_conn,
                  const double *dep_w, const double *hdim, const double *s,
                  const uint8_t *bcmask, double E, double nu) {
  return assemble_impl(nnodes, nelem, conn, dep_ptr, dep_conn, dep_w, hdim, s, 0, bcmask, E, nu);
}

/* Nodal design (TMROctConstitutive with per-Gauss-point penalty,
   src/topology/TMROctConstitutive.cpp:498-551): K_e = sum_q sgp[8 e + q] K_q */
long orc_assemble_gp(int nnodes, int nelem, const int *conn, const int *dep_ptr,
                     const int *dep_conn, const double *dep_w, const double *hdim,
                     const double *sgp, const uint8_t *bcmask, double E, double nu) {
  return assemble_impl(nnodes, nelem, conn, dep_ptr, dep_conn, dep_w, hdim, sgp, 1, bcmask, E,
                       nu);
}

static long assemble_impl(int nnodes, int nelem, const int *conn, const int *dep_ptr,
                          const int *dep_conn, const double *dep_w, const double *hdim,
                          const double *s, int per_gp, const uint8_t *bcmask, double E, double nu) {
  orc_free();
  long cap = 0;
  for (int e = 0; e < nelem; e++) {
    long m = 0;
    for (int a = 0; a < 8; a++) {
      int g = conn[8 * e + a];
      m += (g >= 0) ? 1 : dep_ptr[-g] - dep_ptr[-g - 1];
    }
    cap += 9 * m * m;
  }
  cap += 3L * nnodes;
  trip_t *T = (trip_t *)malloc(sizeof(trip_t) * (size_t)cap);
  long nt = 0;
  double K[576];
  for (int e = 0; e < nelem; e++) {
    double se = 1.0;
    if (per_gp) {
      hex8_stiffness_scaled(&hdim[3 * e], E, nu, &s[8 * (size_t)e], K);
    } else {
      orc_hex8_stiffness(hdim[3 * e], hdim[3 * e + 1], hdim[3 * e + 2], E, nu, K);
      se = s[e];
    }
    int idx[8][4];
    double wt[8][4];
    int cnt[8];
    for (int a = 0; a < 8; a++) {
      int g = conn[8 * e + a];
      if (g >= 0) {
        cnt[a] = 1;
        idx[a][0] = g;
        wt[a][0] = 1.0;
      } else {
        int d = -g - 1;
        cnt[a] = dep_ptr[d + 1] - dep_ptr[d];
        for (int k = 0; k < cnt[a]; k++) {
          idx[a][k] = dep_conn[dep_ptr[d] + k];
          wt[a][k] = dep_w[dep_ptr[d] + k];
        }
      }
    }
    for (int a = 0; a < 8; a++)
      for (int ka = 0; ka < cnt[a]; ka++)
        for (int b = 0; b < 8; b++)
          for (int kb = 0; kb < cnt[b]; kb++) {
            int p = idx[a][ka], q = idx[b][kb];
            double w = se * wt[a][ka] * wt[b][kb];
            for (int i = 0; i < 3; i++) {
              if ((bcmask[p] >> i) & 1) continue;
              for (int j = 0; j < 3; j++) {
                if ((bcmask[q] >> j) & 1) continue;
                T[nt].r = 3 * p + i;
                T[nt].c = 3 * q + j;
                T[nt].v = w * K[24 * (3 * a + i) + 3 * b + j];
                nt++;
              }
            }
          }
  }
  for (int n = 0; n < nnodes; n++)
    for (int i = 0; i < 3; i++)
      if ((bcmask[n] >> i) & 1) {
        T[nt].r = T[nt].c = 3 * n + i;
        T[nt].v = 1.0;
        nt++;
      }
  qsort(T, (size_t)nt, sizeof(trip_t), trip_cmp);
  csr_t *A = (csr_t *)calloc(1, sizeof(csr_t));
  A->n = 3 * nnodes;
  A->rowp = (int *)calloc((size_t)A->n + 1, sizeof(int));
  A->cols = (int *)malloc(sizeof(int) * (size_t)nt);
  A->vals = (double *)malloc(sizeof(double) * (size_t)nt);
  long nz = 0;
  for (long k = 0; k < nt; k++) {
    if (nz > 0 && A->cols[nz - 1] == T[k].c && k > 0 && T[k - 1].r == T[k].r) {
      A->vals[nz - 1] += T[k].v;
    } else {
      A->cols[nz] = T[k].c;
      A->vals[nz] = T[k].v;
      A->rowp[T[k].r + 1]++;
      nz++;
    }
  }
  for (int i = 0; i < A->n; i++) A->rowp[i + 1] += A->rowp[i];
  free(T);
  g_A = A;
  return nz;
}

/* copy the assembled CSR out (sizes from orc_assemble) */
int orc_get_csr(int *rowp, int *cols, double *vals) {
  if (!g_A) return -1;
  memcpy(rowp, g_A->rowp, sizeof(int) * ((size_t)g_A->n + 1));
  memcpy(cols, g_A->cols, sizeof(int) * (size_t)g_A->rowp[g_A->n]);
  memcpy(vals, g_A->vals, sizeof(double) * (size_t)g_A->rowp[g_A->n]);
  return 0;
}

/* threads of the CG loops (OpenMP; 1 by default so that test results do not
   depend on the host) */
static int g_threads = 1;
void orc_set_threads(int n) { g_threads = n > 0 ? n : 1; }
int orc_get_threads(void) { return g_threads; }

void orc_matvec(const double *x, double *y) {
  const csr_t *A = g_A;
#pragma omp parallel for schedule(static) num_threads(g_threads)
  for (int i = 0; i < A->n; i++) {
    double s = 0.0;
    for (int k = A->rowp[i]; k < A->rowp[i + 1]; k++) s += A->vals[k] * x[A->cols[k]];
    y[i] = s;
  }
}

/* Jacobi-preconditioned CG from zero; b must be zero on clamped dofs.
   Returns iterations (negative if not converged). */
int orc_solve(const double *b, double *x, double rtol, int maxit, double *relres) {
  const csr_t *A = g_A;
  const int n = A->n;
  double *r = (double *)malloc(sizeof(double) * n), *z = (double *)malloc(sizeof(double) * n);
  double *p = (double *)malloc(sizeof(double) * n), *q = (double *)malloc(sizeof(double) * n);
  double *dinv = (double *)malloc(sizeof(double) * n);
  for (int i = 0; i < n; i++) {
    double d = 1.0;
    for (int k = A->rowp[i]; k < A->rowp[i + 1]; k++)
      if (A->cols[k] == i) d = A->vals[k];
    dinv[i] = 1.0 / d;
    x[i] = 0.0;
    r[i] = b[i];
  }
  double bb = 0.0;
  for (int i = 0; i < n; i++) bb += b[i] * b[i];
  double bn = sqrt(bb);
  int it = 0, conv = (bn == 0.0);
  double rz = 0.0;
  for (int i = 0; i < n; i++) {
    z[i] = dinv[i] * r[i];
    p[i] = z[i];
    rz += r[i] * z[i];
  }
  double rn = bn;
  while (!conv && it < maxit) {
    it++;
    orc_matvec(p, q);
    double pq = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : pq) num_threads(g_threads)
    for (int i = 0; i < n; i++) pq += p[i] * q[i];
    double alpha = rz / pq;
    double rr = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : rr) num_threads(g_threads)
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
    double rzn = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : rzn) num_threads(g_threads)
    for (int i = 0; i < n; i++) {
      z[i] = dinv[i] * r[i];
      rzn += r[i] * z[i];
    }
    double beta = rzn / rz;
    rz = rzn;
#pragma omp parallel for schedule(static) num_threads(g_threads)
    for (int i = 0; i < n; i++) p[i] = z[i] + beta * p[i];
  }
  /* true residual */
  orc_matvec(x, q);
  double tr = 0.0;
  for (int i = 0; i < n; i++) tr += (b[i] - q[i]) * (b[i] - q[i]);
  if (relres) *relres = bn > 0 ? sqrt(tr) / bn : 0.0;
  free(r);
  free(z);
  free(p);
  free(q);
  free(dinv);
  return conv ? it : -it;
}

/* Element stiffness of box extents h, memoised: a forest has few distinct
   element sizes (levels x block shapes).  Returns NULL past 256 sizes. */
static const double *k_cached(const double h[3], double E, double nu) {
  static double keys[256][5];
  static double *vals[256];
  static int nk = 0;
  for (int i = 0; i < nk; i++)
    if (keys[i][0] == h[0] && keys[i][1] == h[1] && keys[i][2] == h[2] && keys[i][3] == E &&
        keys[i][4] == nu)
      return vals[i];
  if (nk == 256) return NULL;
  double *K = (double *)malloc(sizeof(double) * 576);
  orc_hex8_stiffness(h[0], h[1], h[2], E, nu, K);
  keys[nk][0] = h[0];
  keys[nk][1] = h[1];
  keys[nk][2] = h[2];
  keys[nk][3] = E;
  keys[nk][4] = nu;
  vals[nk] = K;
  return vals[nk++];
}

/* element energies u_e^T K_e u_e / s_e (unpenalised) */
void orc_elem_energy(int nelem, const int *conn, const int *dep_ptr, const int *dep_conn,
                     const double *dep_w, const double *hdim, const uint8_t *bcmask, double E,
                     double nu, const double *u, double *energy) {
  double Kbuf[576], ue[24];
  for (int e = 0; e < nelem; e++) {
    const double *K = k_cached(&hdim[3 * e], E, nu);
    if (!K) {
      orc_hex8_stiffness(hdim[3 * e], hdim[3 * e + 1], hdim[3 * e + 2], E, nu, Kbuf);
      K = Kbuf;
    }
    for (int a = 0; a < 8; a++) {
      int g = conn[8 * e + a];
      for (int i = 0; i < 3; i++) {
        double v = 0.0;
        if (g >= 0) {
          v = ((bcmask[g] >> i) & 1) ? 0.0 : u[3 * g + i];
        } else {
          int d = -g - 1;
          for (int k = dep_ptr[d]; k < dep_ptr[d + 1]; k++) {
            int p = dep_conn[k];
            if (!((bcmask[p] >> i) & 1)) v += dep_w[k] * u[3 * p + i];
          }
        }
        ue[3 * a + i] = v;
      }
    }
    double s = 0.0;
    for (int i = 0; i < 24; i++)
      for (int j = 0; j < 24; j++) s += ue[i] * K[24 * i + j] * ue[j];
    energy[e] = s;
  }
}

/* ---- p-norm of the relaxed von Mises failure ----------------------------
   f = (sum_e sum_g (V_e/8) fail_eg^P)^(1/P), fail = r(rho) vm(C e_g) / ys,
   r = rho / (eps (1 - rho) + rho) (TMROctConstitutive::evalFailure,
   src/topology/TMROctConstitutive.cpp:1056-1104; aggregation as TACS's
   "pnorm-continuous" KSFailure).  Returns f; when dfdu / dfdrho are not
   NULL they receive df/du (assembled through the hanging-node weights,
   clamped dofs zero) and the explicit df/drho_e. */
static void gather_ue(const int *conn, int e, const int *dep_ptr, const int *dep_conn,
                      const double *dep_w, const uint8_t *bcmask, const double *u, double ue[24]) {
  for (int a = 0; a < 8; a++) {
    int g = conn[8 * e + a];
    for (int i = 0; i < 3; i++) {
      double v = 0.0;
      if (g >= 0) {
        v = ((bcmask[g] >> i) & 1) ? 0.0 : u[3 * g + i];
      } else {
        int d = -g - 1;
        for (int k = dep_ptr[d]; k < dep_ptr[d + 1]; k++) {
          int p = dep_conn[k];
          if (!((bcmask[p] >> i) & 1)) v += dep_w[k] * u[3 * p + i];
        }
      }
      ue[3 * a + i] = v;
    }
  }
}

static double stress_impl(int nnodes, int nelem, const int *conn, const int *dep_ptr,
                          const int *dep_conn, const double *dep_w, const double *hdim,
                          const uint8_t *bcmask, double E, double nu, double ys, double eps,
                          double P, const double *rho, int per_gp, const double *u, double *dfdu,
                          double *dfdrho);

double orc_stress_pnorm(int nnodes, int nelem, const int *conn, const int *dep_ptr,
                        const int *dep_conn, const double *dep_w, const double *hdim,
                        const uint8_t *bcmask, double E, double nu, double ys, double eps,
                        double P, const double *rho, const double *u, double *dfdu,
                        double *dfdrho) {
  return stress_impl(nnodes, nelem, conn, dep_ptr, dep_conn, dep_w, hdim, bcmask, E, nu, ys, eps,
                     P, rho, 0, u, dfdu, dfdrho);
}

/* nodal design: rho[8 e + q] at the Gauss points, dfdrho per Gauss point */
double orc_stress_pnorm_gp(int nnodes, int nelem, const int *conn, const int *dep_ptr,
                           const int *dep_conn, const double *dep_w, const double *hdim,
                           const uint8_t *bcmask, double E, double nu, double ys, double eps,
                           double P, const double *rho_gp, const double *u, double *dfdu,
                           double *dfdrho_gp) {
  return stress_impl(nnodes, nelem, conn, dep_ptr, dep_conn, dep_w, hdim, bcmask, E, nu, ys, eps,
                     P, rho_gp, 1, u, dfdu, dfdrho_gp);
}

static double stress_impl(int nnodes, int nelem, const int *conn, const int *dep_ptr,
                          const int *dep_conn, const double *dep_w, const double *hdim,
                          const uint8_t *bcmask, double E, double nu, double ys, double eps,
                          double P, const double *rho, int per_gp, const double *u, double *dfdu,
                          double *dfdrho) {
  double C[6][6];
  iso_C(E, nu, C);
  const double gp = 1.0 / sqrt(3.0);
  double S = 0.0;
  /* first pass: the sum */
  for (int e = 0; e < nelem; e++) {
    double ue[24], h[3] = {hdim[3 * e], hdim[3 * e + 1], hdim[3 * e + 2]};
    gather_ue(conn, e, dep_ptr, dep_conn, dep_w, bcmask, u, ue);
    double V = h[0] * h[1] * h[2];
    for (int q = 0; q < 8; q++) {
      double rq = per_gp ? rho[8 * (size_t)e + q] : rho[e];
      double r = rq / (eps * (1.0 - rq) + rq);
      double p[3] = {(q & 1) ? gp : -gp, (q & 2) ? gp : -gp, (q & 4) ? gp : -gp};
      double ep[6], s[6];
      strain_at(h, p, ue, ep);
      for (int i = 0; i < 6; i++) {
        s[i] = 0.0;
        for (int j = 0; j < 6; j++) s[i] += C[i][j] * ep[j];
      }
      double vm = sqrt(0.5 * ((s[0] - s[1]) * (s[0] - s[1]) + (s[1] - s[2]) * (s[1] - s[2]) +
                              (s[2] - s[0]) * (s[2] - s[0])) +
                       3.0 * (s[3] * s[3] + s[4] * s[4] + s[5] * s[5])) / ys;
      S += 0.125 * V * pow(r * vm, P);
    }
  }
  double f = pow(S, 1.0 / P);
  if (!dfdu && !dfdrho) return f;
  double coef = pow(S, 1.0 / P - 1.0);
  if (dfdu) memset(dfdu, 0, sizeof(double) * 3 * (size_t)nnodes);
  for (int e = 0; e < nelem; e++) {
    double ue[24], h[3] = {hdim[3 * e], hdim[3 * e + 1], hdim[3 * e + 2]};
    gather_ue(conn, e, dep_ptr, dep_conn, dep_w, bcmask, u, ue);
    double V = h[0] * h[1] * h[2];
    double ge[24] = {0}, sumvmP = 0.0;
    for (int q = 0; q < 8; q++) {
      double rq = per_gp ? rho[8 * (size_t)e + q] : rho[e];
      double den = eps * (1.0 - rq) + rq;
      double r = rq / den, dr = eps / (den * den);
      double p[3] = {(q & 1) ? gp : -gp, (q & 2) ? gp : -gp, (q & 4) ? gp : -gp};
      double ep[6], s[6];
      strain_at(h, p, ue, ep);
      for (int i = 0; i < 6; i++) {
        s[i] = 0.0;
        for (int j = 0; j < 6; j++) s[i] += C[i][j] * ep[j];
      }
      double J = 0.5 * ((s[0] - s[1]) * (s[0] - s[1]) + (s[1] - s[2]) * (s[1] - s[2]) +
                        (s[2] - s[0]) * (s[2] - s[0])) +
                 3.0 * (s[3] * s[3] + s[4] * s[4] + s[5] * s[5]);
      double vm = sqrt(J) / ys;
      sumvmP += pow(vm, P);
      if (per_gp && dfdrho) dfdrho[8 * (size_t)e + q] = coef * 0.125 * V * pow(r, P - 1.0) * dr * pow(vm, P);
      /* d fail^P / d s = P fail^(P-1) r dvm/ds, dvm/ds = dJ/ds / (2 ys^2 vm) */
      double dJ[6] = {2 * s[0] - s[1] - s[2], 2 * s[1] - s[0] - s[2], 2 * s[2] - s[0] - s[1],
                      6 * s[3], 6 * s[4], 6 * s[5]};
      double w = coef * 0.125 * V * pow(r, P) * pow(vm, P - 2.0) / (2.0 * ys * ys);
      /* chain through s = C eps(u): dfdu_c = sum_i w dJ_i (C eps(e_c))_i */
      double unit[24];
      for (int c = 0; c < 24; c++) {
        memset(unit, 0, sizeof(unit));
        unit[c] = 1.0;
        double ec[6];
        strain_at(h, p, unit, ec);
        double t = 0.0;
        for (int i = 0; i < 6; i++) {
          double sc = 0.0;
          for (int j = 0; j < 6; j++) sc += C[i][j] * ec[j];
          t += dJ[i] * sc;
        }
        ge[c] += w * t;
      }
    }
    if (!per_gp && dfdrho) {
      double den = eps * (1.0 - rho[e]) + rho[e];
      double r = rho[e] / den, dr = eps / (den * den);
      dfdrho[e] = coef * 0.125 * V * pow(r, P - 1.0) * dr * sumvmP;
    }
    if (dfdu) {
      for (int a = 0; a < 8; a++) {
        int g = conn[8 * e + a];
        for (int i = 0; i < 3; i++) {
          if (g >= 0) {
            if (!((bcmask[g] >> i) & 1)) dfdu[3 * g + i] += ge[3 * a + i];
          } else {
            int d = -g - 1;
            for (int k = dep_ptr[d]; k < dep_ptr[d + 1]; k++) {
              int p = dep_conn[k];
              if (!((bcmask[p] >> i) & 1)) dfdu[3 * p + i] += dep_w[k] * ge[3 * a + i];
            }
          }
        }
      }
    }
  }
  return f;
}

/* psi_e^T K_e u_e per element (unpenalised) */
void orc_elem_energy2(int nelem, const int *conn, const int *dep_ptr, const int *dep_conn,
                      const double *dep_w, const double *hdim, const uint8_t *bcmask, double E,
                      double nu, const double *u, const double *v, double *out) {
  double K[576], ue[24], ve[24];
  for (int e = 0; e < nelem; e++) {
    orc_hex8_stiffness(hdim[3 * e], hdim[3 * e + 1], hdim[3 * e + 2], E, nu, K);
    gather_ue(conn, e, dep_ptr, dep_conn, dep_w, bcmask, u, ue);
    gather_ue(conn, e, dep_ptr, dep_conn, dep_w, bcmask, v, ve);
    double s = 0.0;
    for (int i = 0; i < 24; i++)
      for (int j = 0; j < 24; j++) s += ve[i] * K[24 * i + j] * ue[j];
    out[e] = s;
  }
}

/* v_e^T K_q u_e at every Gauss point q of every element (unpenalised),
   out[8 e + q] */
void orc_elem_energy_gp(int nelem, const int *conn, const int *dep_ptr, const int *dep_conn,
                        const double *dep_w, const double *hdim, const uint8_t *bcmask, double E,
                        double nu, const double *u, const double *v, double *out) {
  double K[576], ue[24], ve[24];
  for (int e = 0; e < nelem; e++) {
    gather_ue(conn, e, dep_ptr, dep_conn, dep_w, bcmask, u, ue);
    gather_ue(conn, e, dep_ptr, dep_conn, dep_w, bcmask, v, ve);
    for (int q = 0; q < 8; q++) {
      orc_hex8_stiffness_gp(hdim[3 * e], hdim[3 * e + 1], hdim[3 * e + 2], E, nu, q, K);
      double t = 0.0;
      for (int i = 0; i < 24; i++)
        for (int j = 0; j < 24; j++) t += ve[i] * K[24 * i + j] * ue[j];
      out[8 * (size_t)e + q] = t;
    }
  }
}

/* Matrix-free y = A x with A = Z^T (sum_e s_e K_e) Z, identity rows and
   columns on clamped dofs (the operator orc_assemble stores), element by
   element without assembly: the residual check of solutions too large to
   assemble.  s has one factor per element. */
void orc_apply_free(int nnodes, int nelem, const int *conn, const int *dep_ptr,
                    const int *dep_conn, const double *dep_w, const double *hdim, const double *s,
                    const uint8_t *bcmask, double E, double nu, const double *x, double *y) {
  double Kbuf[576], ue[24], ye[24];
  memset(y, 0, sizeof(double) * 3 * (size_t)nnodes);
  for (int e = 0; e < nelem; e++) {
    const double *K = k_cached(&hdim[3 * e], E, nu);
    if (!K) {
      orc_hex8_stiffness(hdim[3 * e], hdim[3 * e + 1], hdim[3 * e + 2], E, nu, Kbuf);
      K = Kbuf;
    }
    gather_ue(conn, e, dep_ptr, dep_conn, dep_w, bcmask, x, ue);
    for (int i = 0; i < 24; i++) {
      double t = 0.0;
      for (int j = 0; j < 24; j++) t += K[24 * i + j] * ue[j];
      ye[i] = s[e] * t;
    }
    for (int a = 0; a < 8; a++) {
      const int g = conn[8 * e + a];
      for (int i = 0; i < 3; i++) {
        if (g >= 0) {
          if (!((bcmask[g] >> i) & 1)) y[3 * (size_t)g + i] += ye[3 * a + i];
        } else {
          const int d = -g - 1;
          for (int k = dep_ptr[d]; k < dep_ptr[d + 1]; k++) {
            const int p = dep_conn[k];
            if (!((bcmask[p] >> i) & 1)) y[3 * (size_t)p + i] += dep_w[k] * ye[3 * a + i];
          }
        }
      }
    }
  }
  for (int n = 0; n < nnodes; n++)
    for (int i = 0; i < 3; i++)
      if ((bcmask[n] >> i) & 1) y[3 * (size_t)n + i] = x[3 * (size_t)n + i];
}
