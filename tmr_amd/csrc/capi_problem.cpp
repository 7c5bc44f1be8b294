// C-ABI wrappers for the solver (declarations in include/tmrx.h).
#include <hip/hip_runtime_api.h>

#include <string>
#include <vector>

#include "../../include/tmrx.h"
#include "problem.h"

namespace tmrx {
void set_error(const std::string &s);
}
using tmrx::set_error;
tmrx::Forest *tmrx_forest_impl(tmrx_forest *f);

struct tmrx_problem {
  tmrx::Problem *p;
  int device;
};

static tmrx::SolveOpts to_opts(const tmrx_solve_opts *o) {
  tmrx::SolveOpts s;
  if (!o) return s;
  s.rtol = o->rtol;
  s.atol = o->atol;
  s.max_iters = o->max_iters;
  s.precision = o->precision;
  s.cheb_degree = o->cheb_degree;
  s.cheb_lower = o->cheb_lower;
  s.cheb_upper = o->cheb_upper;
  s.eig_iters = o->eig_iters;
  s.monitor = o->monitor;
  s.zero_guess = o->zero_guess;
  s.fp32_valu = o->fp32_valu;
  s.no_graph = o->no_graph;
  s.cheb_pre_degree = o->cheb_pre_degree;
  return s;
}

#define CHECK_P(h)                         \
  do {                                     \
    if (!(h) || !(h)->p) {                 \
      set_error("null problem handle");    \
      return -1;                           \
    }                                      \
    (void)hipSetDevice((h)->device);       \
  } while (0)

static int fail(tmrx_problem *h) {
  set_error(h->p->last_error);
  return -1;
}

extern "C" {

void tmrx_solve_opts_default(tmrx_solve_opts *o) {
  if (!o) return;
  tmrx::SolveOpts s;
  o->rtol = s.rtol;
  o->atol = s.atol;
  o->max_iters = s.max_iters;
  o->precision = s.precision;
  o->cheb_degree = s.cheb_degree;
  o->cheb_lower = s.cheb_lower;
  o->cheb_upper = s.cheb_upper;
  o->eig_iters = s.eig_iters;
  o->monitor = s.monitor;
  o->zero_guess = s.zero_guess;
  o->fp32_valu = s.fp32_valu;
  o->no_graph = s.no_graph;
  o->cheb_pre_degree = s.cheb_pre_degree;
}

int tmrx_problem_create(int nlevels, tmrx_forest **forests, const tmrx_material *mat, int nbcs,
                        const char **bc_names, const int *bc_dof_masks, int device,
                        tmrx_problem **out) {
  if (!out || !forests || nlevels < 1 || nbcs < 0 || (nbcs > 0 && !bc_names)) {
    set_error("tmrx_problem_create: invalid arguments");
    return -1;
  }
  // host-side checks first (no device needed to find these mistakes)
  if (mat && (!(mat->E > 0.0) || !(mat->nu > -1.0 && mat->nu < 0.5) || !(mat->k0 >= 0.0) ||
              !(mat->q >= 0.0) || !(mat->ys > 0.0) || !(mat->eps >= 0.0) ||
              (mat->penalty_type != 0 && mat->penalty_type != 1))) {
    set_error("tmrx_problem_create: material needs E > 0, -1 < nu < 0.5, ys > 0, q >= 0, "
              "k0 >= 0, eps >= 0 and penalty_type 0 (SIMP) or 1 (RAMP)");
    return -1;
  }
  {
    tmrx::Forest *f0 = tmrx_forest_impl(forests[0]);
    if (!f0) {
      set_error("tmrx_problem_create: null forest");
      return -1;
    }
    for (int k = 0; k < nbcs; k++) {
      if (!bc_names[k]) {
        set_error("tmrx_problem_create: null boundary-condition name");
        return -1;
      }
      if (f0->create_nodes()) {
        set_error("tmrx_problem_create: " + f0->last_error);
        return -1;
      }
      std::vector<int> nodes;
      f0->nodes_with_name(bc_names[k], nodes);
      if (nodes.empty()) {
        set_error(std::string("tmrx_problem_create: boundary condition '") + bc_names[k] +
                  "' selects no node of the finest forest");
        return -1;
      }
    }
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
    set_error("tmrx_problem_create: no HIP device available");
    return -1;
  }
  if (device < 0 || device >= ndev) {
    set_error("tmrx_problem_create: device index out of range");
    return -1;
  }
  if (hipSetDevice(device) != hipSuccess) {
    set_error("tmrx_problem_create: hipSetDevice failed");
    return -1;
  }
  std::vector<tmrx::Forest *> fs(nlevels);
  for (int l = 0; l < nlevels; l++) {
    fs[l] = tmrx_forest_impl(forests[l]);
    if (!fs[l]) {
      set_error("tmrx_problem_create: null forest");
      return -1;
    }
  }
  tmrx::Material m;
  if (mat) {
    m.E = mat->E;
    m.nu = mat->nu;
    m.ys = mat->ys;
    m.q = mat->q;
    m.k0 = mat->k0;
    m.eps = mat->eps;
    m.penalty_type = mat->penalty_type;
  }
  if (m.penalty_type == 0 && m.q < 1.0) m.q = 1.0;  // TMRStiffnessProperties clamp
  auto *p = new tmrx::Problem();
  if (p->init(nlevels, fs.data(), m, nbcs, bc_names, bc_dof_masks)) {
    set_error(p->last_error);
    delete p;
    return -1;
  }
  *out = new tmrx_problem{p, device};
  return 0;
}

void tmrx_problem_free(tmrx_problem *h) {
  if (!h) return;
  (void)hipSetDevice(h->device);
  delete h->p;
  delete h;
}

int tmrx_problem_get_sizes(tmrx_problem *h, int *ndof, int *nelem, int *nlevels) {
  CHECK_P(h);
  if (ndof) *ndof = h->p->num_dof();
  if (nelem) *nelem = h->p->num_elements();
  if (nlevels) *nlevels = h->p->num_levels();
  return 0;
}

int tmrx_problem_set_stream(tmrx_problem *h, void *stream) {
  CHECK_P(h);
  h->p->set_stream((hipStream_t)stream);
  return 0;
}

int tmrx_problem_set_design(tmrx_problem *h, const double *rho, int n, int on_device) {
  CHECK_P(h);
  if (!rho) {
    set_error("tmrx_problem_set_design: null densities");
    return -1;
  }
  return h->p->set_design(rho, n, on_device != 0) ? fail(h) : 0;
}

int tmrx_problem_set_design_nodal(tmrx_problem *h, const double *x, int n, int on_device) {
  CHECK_P(h);
  if (!x) {
    set_error("tmrx_problem_set_design_nodal: null design vector");
    return -1;
  }
  return h->p->set_design_nodal(x, n, on_device != 0) ? fail(h) : 0;
}

int tmrx_problem_set_bc_values(tmrx_problem *h, const double *ubc, int n, int on_device) {
  CHECK_P(h);
  if (!ubc) {
    set_error("tmrx_problem_set_bc_values: null values");
    return -1;
  }
  return h->p->set_bc_values(ubc, n, on_device != 0) ? fail(h) : 0;
}

int tmrx_problem_design_info(tmrx_problem *h, int *mode, int *n) {
  CHECK_P(h);
  if (mode) *mode = h->p->design_mode();
  if (n) *n = h->p->design_size();
  return 0;
}

int tmrx_problem_level_design(tmrx_problem *h, int level, double *x) {
  CHECK_P(h);
  if (!x) {
    set_error("tmrx_problem_level_design: null output");
    return -1;
  }
  return h->p->level_design(level, x) ? fail(h) : 0;
}

int tmrx_problem_setup(tmrx_problem *h, const tmrx_solve_opts *opts) {
  CHECK_P(h);
  return h->p->setup(to_opts(opts)) ? fail(h) : 0;
}

int tmrx_problem_solve(tmrx_problem *h, const double *f, double *u, int on_device,
                       const tmrx_solve_opts *opts, tmrx_solve_info *info) {
  CHECK_P(h);
  tmrx::SolveInfo si;
  int r = h->p->solve(f, u, on_device != 0, to_opts(opts), &si);
  if (info) {
    info->iters = si.iters;
    info->converged = si.converged;
    info->rnorm0 = si.rnorm0;
    info->rnorm = si.rnorm;
    info->setup_ms = si.setup_ms;
    info->solve_ms = si.solve_ms;
  }
  if (r < 0) return fail(h);
  if (r > 0) set_error(h->p->last_error);
  return r;
}

int tmrx_problem_solve_multi(tmrx_problem *h, int nrhs, const double *f, double *u, int on_device,
                             const tmrx_solve_opts *opts, tmrx_solve_info *infos) {
  CHECK_P(h);
  if (nrhs < 1) {
    set_error("tmrx_problem_solve_multi: nrhs must be positive");
    return -1;
  }
  std::vector<tmrx::SolveInfo> si(nrhs);
  int r = h->p->solve_multi(nrhs, f, u, on_device != 0, to_opts(opts), si.data());
  if (infos)
    for (int k = 0; k < nrhs; k++) {
      infos[k].iters = si[k].iters;
      infos[k].converged = si[k].converged;
      infos[k].rnorm0 = si[k].rnorm0;
      infos[k].rnorm = si[k].rnorm;
      infos[k].setup_ms = si[k].setup_ms;
      infos[k].solve_ms = si[k].solve_ms;
    }
  if (r < 0) return fail(h);
  if (r > 0) set_error(h->p->last_error);
  return r;
}

int tmrx_problem_autotune(tmrx_problem *h, const int *tile_sizes, int n, int reps, int *chosen,
                          double *ms) {
  CHECK_P(h);
  return h->p->autotune(tile_sizes, n, reps, chosen, ms) ? fail(h) : 0;
}

int tmrx_problem_tuning(tmrx_problem *h, int *tile_nodes, int *fp32_matrix_cores) {
  CHECK_P(h);
  if (tile_nodes) *tile_nodes = h->p->tile_nodes();
  if (fp32_matrix_cores) *fp32_matrix_cores = h->p->fp32_form();
  return 0;
}

int tmrx_problem_tile_info(tmrx_problem *h, int level, int *ntiles, int *tile_nodes) {
  CHECK_P(h);
  if (h->p->tile_info(level, ntiles, tile_nodes)) {
    set_error("tmrx_problem_tile_info: no such level");
    return -1;
  }
  return 0;
}

int tmrx_problem_set_tiling(tmrx_problem *h, int tile_nodes) {
  CHECK_P(h);
  return h->p->set_tiling(tile_nodes) ? fail(h) : 0;
}

int tmrx_problem_set_fp32_form(tmrx_problem *h, int form) {
  CHECK_P(h);
  return h->p->set_fp32_form(form) ? fail(h) : 0;
}

int tmrx_mg_create(tmrx_problem *h, double omega, int use_galerkin, int use_coarse_direct_solve,
                   int use_chebyshev_smoother) {
  CHECK_P(h);
  if (!(omega > 0.0 && omega < 2.0)) {
    set_error("tmrx_mg_create: omega must lie in (0, 2)");
    return -1;
  }
  if (h->p->set_galerkin(use_galerkin != 0)) return fail(h);
  if (h->p->set_coarse_direct(use_coarse_direct_solve != 0)) return fail(h);
  if (h->p->set_smoother(use_chebyshev_smoother == 0, omega)) return fail(h);
  return 0;
}

int tmrx_problem_smoother_order(tmrx_problem *h, int level, int *nnodes, int *ncolors, int *nodes,
                                int *color_ptr) {
  CHECK_P(h);
  std::vector<int> nd, cp;
  if (h->p->smoother_order(level, nd, cp)) return fail(h);
  if (nnodes) *nnodes = (int)nd.size();
  if (ncolors) *ncolors = (int)cp.size() - 1;
  if (nodes) std::copy(nd.begin(), nd.end(), nodes);
  if (color_ptr) std::copy(cp.begin(), cp.end(), color_ptr);
  return 0;
}

int tmrx_problem_level_matrix(tmrx_problem *h, int level, int *n, int *nnzb, int *ptr, int *col,
                              double *val) {
  CHECK_P(h);
  std::vector<int> p, c;
  std::vector<double> v;
  if (h->p->level_matrix(level, p, c, v)) return fail(h);
  if (n) *n = (int)p.size() - 1;
  if (nnzb) *nnzb = (int)c.size();
  if (ptr) std::copy(p.begin(), p.end(), ptr);
  if (col) std::copy(c.begin(), c.end(), col);
  if (val) std::copy(v.begin(), v.end(), val);
  return 0;
}

int tmrx_problem_stress_pnorm_noalloc(tmrx_problem *h, double P, double *value) {
  CHECK_P(h);
  if (!value) {
    set_error("tmrx_problem_stress_pnorm_noalloc: null output");
    return -1;
  }
  if (h->p->stress_pnorm_noalloc(P, value)) return fail(h);
  return 0;
}

int tmrx_problem_num_components(tmrx_problem *h, int *ncomp) {
  CHECK_P(h);
  if (ncomp) *ncomp = h->p->num_segments();
  return 0;
}

int tmrx_problem_component_info(tmrx_problem *h, int rhs, int comp, int *iters, int *converged,
                                double *rel_residual) {
  CHECK_P(h);
  if (h->p->segment_info(rhs, comp, iters, converged, rel_residual)) {
    set_error("tmrx_problem_component_info: no such right-hand side / component in the last solve");
    return -1;
  }
  return 0;
}

int tmrx_problem_true_residual(tmrx_problem *h, double *rel) {
  CHECK_P(h);
  if (rel) *rel = h->p->last_true_residual;
  return 0;
}

int tmrx_problem_set_state(tmrx_problem *h, const double *f, const double *u, int on_device) {
  CHECK_P(h);
  return h->p->set_state(f, u, on_device != 0) ? fail(h) : 0;
}

int tmrx_problem_compliance(tmrx_problem *h, double *c) {
  CHECK_P(h);
  return h->p->compliance(c) ? fail(h) : 0;
}

int tmrx_problem_compliance_grad(tmrx_problem *h, double *g, int on_device) {
  CHECK_P(h);
  return h->p->compliance_grad(g, on_device != 0) ? fail(h) : 0;
}

int tmrx_problem_mg_op(tmrx_problem *h, int op, int level, const double *x, double *y,
                       int on_device, const tmrx_solve_opts *opts, double *lam_max) {
  CHECK_P(h);
  return h->p->mg_op(op, level, x, y, on_device != 0, to_opts(opts), lam_max) ? fail(h) : 0;
}

int tmrx_problem_apply(tmrx_problem *h, const double *x, double *y, int on_device,
                       int precision) {
  CHECK_P(h);
  return h->p->apply(x, y, on_device != 0, precision) ? fail(h) : 0;
}

int tmrx_problem_traction_load(tmrx_problem *h, const char *name, const double tr[3],
                               double *f) {
  CHECK_P(h);
  if (!name || !tr || !f) {
    set_error("tmrx_problem_traction_load: invalid arguments");
    return -1;
  }
  return h->p->traction_load(name, tr, f) ? fail(h) : 0;
}

int tmrx_problem_stress_pnorm(tmrx_problem *h, double P, double *value) {
  CHECK_P(h);
  if (!value) {
    set_error("tmrx_problem_stress_pnorm: null output");
    return -1;
  }
  return h->p->stress_pnorm(P, value) ? fail(h) : 0;
}

int tmrx_problem_stress_pnorm_grad(tmrx_problem *h, double P, double *g, int on_device,
                                   const tmrx_solve_opts *opts, tmrx_solve_info *info) {
  CHECK_P(h);
  if (!g) {
    set_error("tmrx_problem_stress_pnorm_grad: null output");
    return -1;
  }
  tmrx::SolveInfo si;
  int r = h->p->stress_pnorm_grad(P, g, on_device != 0, to_opts(opts), &si);
  if (info) {
    info->iters = si.iters;
    info->converged = si.converged;
    info->rnorm0 = si.rnorm0;
    info->rnorm = si.rnorm;
    info->setup_ms = si.setup_ms;
    info->solve_ms = si.solve_ms;
  }
  if (r < 0) return fail(h);
  if (r > 0) set_error(h->p->last_error);
  return r;
}

int tmrx_problem_kernel_timing(tmrx_problem *h, int enable) {
  CHECK_P(h);
  return h->p->kernel_timing(enable);
}

int tmrx_problem_kernel_timing_result(tmrx_problem *h, double *total_ms, int *count,
                                      int *precision_mask) {
  CHECK_P(h);
  return h->p->kernel_timing_result(total_ms, count, precision_mask) ? fail(h) : 0;
}

}  // extern "C"
