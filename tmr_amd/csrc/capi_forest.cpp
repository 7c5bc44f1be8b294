// C-ABI wrappers for the forest (declarations in include/tmrx.h).
#include <cstring>
#include <string>

#include "../../include/tmrx.h"
#include "forest.h"

struct tmrx_forest {
  tmrx::Forest *f;
  tmrx::InterpCSR P;
};

namespace tmrx {
thread_local std::string g_last_error;
void set_error(const std::string &s) { g_last_error = s; }
}  // namespace tmrx

using tmrx::set_error;

static int fail_from(tmrx_forest *h) {
  set_error(h->f->last_error);
  return -1;
}

#define CHECK_HANDLE(h)                         \
  do {                                          \
    if (!(h) || !(h)->f) {                      \
      set_error("null forest handle");          \
      return -1;                                \
    }                                           \
  } while (0)

extern "C" {

const char *tmrx_last_error(void) { return tmrx::g_last_error.c_str(); }
const char *tmrx_version(void) { return "tmrx 0.1 (gfx950)"; }

int tmrx_forest_create(int mesh_order, int interp_type, tmrx_forest **out) {
  if (!out) {
    set_error("tmrx_forest_create: null output");
    return -1;
  }
  if (mesh_order < 2) mesh_order = 2;
  *out = new tmrx_forest{new tmrx::Forest(mesh_order, interp_type), {}};
  return 0;
}

void tmrx_forest_free(tmrx_forest *f) {
  if (!f) return;
  delete f->f;
  delete f;
}

int tmrx_forest_set_connectivity(tmrx_forest *f, int num_nodes, const int *block_conn,
                                 int num_blocks) {
  CHECK_HANDLE(f);
  return f->f->set_connectivity(num_nodes, block_conn, num_blocks) ? fail_from(f) : 0;
}

int tmrx_forest_get_connectivity(tmrx_forest *f, int *nblocks, int *nfaces, int *nedges,
                                 int *nnodes, const int **block_conn,
                                 const int **block_face_conn, const int **block_edge_conn,
                                 const int **block_face_ids) {
  CHECK_HANDLE(f);
  const tmrx::BlockConn &b = f->f->conn_data();
  if (nblocks) *nblocks = b.num_blocks;
  if (nfaces) *nfaces = b.num_faces;
  if (nedges) *nedges = b.num_edges;
  if (nnodes) *nnodes = b.num_nodes;
  if (block_conn) *block_conn = b.block_conn.data();
  if (block_face_conn) *block_face_conn = b.block_face_conn.data();
  if (block_edge_conn) *block_edge_conn = b.block_edge_conn.data();
  if (block_face_ids) *block_face_ids = b.block_face_ids.data();
  return 0;
}

int tmrx_forest_set_node_points(tmrx_forest *f, const double *Xpts) {
  CHECK_HANDLE(f);
  if (!Xpts) {
    set_error("tmrx_forest_set_node_points: null points");
    return -1;
  }
  return f->f->set_node_points(Xpts) ? fail_from(f) : 0;
}

int tmrx_forest_set_name(tmrx_forest *f, int kind, int index, const char *name) {
  CHECK_HANDLE(f);
  return f->f->set_name(kind, index, name) ? fail_from(f) : 0;
}

int tmrx_forest_create_trees(tmrx_forest *f, int level) {
  CHECK_HANDLE(f);
  return f->f->create_trees(level) ? fail_from(f) : 0;
}

int tmrx_forest_create_trees_per_block(tmrx_forest *f, const int *block_levels, int nblocks) {
  CHECK_HANDLE(f);
  if (!block_levels || nblocks != f->f->conn_data().num_blocks) {
    set_error("tmrx_forest_create_trees_per_block: need one level per block");
    return -1;
  }
  return f->f->create_trees(0, block_levels) ? fail_from(f) : 0;
}

int tmrx_forest_set_octants(tmrx_forest *f, const tmrx_octant *octs, int n) {
  CHECK_HANDLE(f);
  return f->f->set_octants(reinterpret_cast<const tmrx::Oct *>(octs), n) ? fail_from(f) : 0;
}

int tmrx_forest_refine(tmrx_forest *f, const int *refinement, int min_level, int max_level) {
  CHECK_HANDLE(f);
  return f->f->refine(refinement, min_level, max_level) ? fail_from(f) : 0;
}

int tmrx_forest_balance(tmrx_forest *f, int balance_corner) {
  CHECK_HANDLE(f);
  return f->f->balance(balance_corner) ? fail_from(f) : 0;
}

int tmrx_forest_balance_device(tmrx_forest *f, int balance_corner, int device, double *ms) {
  CHECK_HANDLE(f);
  return f->f->balance_device(balance_corner, device, ms) ? fail_from(f) : 0;
}

int tmrx_forest_coarsen(tmrx_forest *f, tmrx_forest **out) {
  CHECK_HANDLE(f);
  if (!out) {
    set_error("tmrx_forest_coarsen: null output");
    return -1;
  }
  *out = new tmrx_forest{f->f->coarsen(), {}};
  return 0;
}

int tmrx_forest_duplicate(tmrx_forest *f, tmrx_forest **out) {
  CHECK_HANDLE(f);
  if (!out) {
    set_error("tmrx_forest_duplicate: null output");
    return -1;
  }
  *out = new tmrx_forest{f->f->duplicate(), {}};
  return 0;
}

int tmrx_forest_repartition(tmrx_forest *f) {
  CHECK_HANDLE(f);
  return 0;
}

int tmrx_forest_create_nodes(tmrx_forest *f) {
  CHECK_HANDLE(f);
  return f->f->create_nodes() ? fail_from(f) : 0;
}

int tmrx_forest_create_nodes_device(tmrx_forest *f, int device, double *ms) {
  CHECK_HANDLE(f);
  return f->f->create_nodes_device(device, ms) ? fail_from(f) : 0;
}

int tmrx_forest_get_octants(tmrx_forest *f, const tmrx_octant **octs, int *n) {
  CHECK_HANDLE(f);
  const auto &o = f->f->octants();
  if (octs) *octs = reinterpret_cast<const tmrx_octant *>(o.data());
  if (n) *n = (int)o.size();
  return 0;
}

static int need_nodes(tmrx_forest *f, const char *who) {
  if (!f->f->has_nodes()) {
    set_error(std::string(who) + ": nodes have not been created");
    return -1;
  }
  return 0;
}

int tmrx_forest_get_node_conn(tmrx_forest *f, const int **conn, int *num_elements,
                              int *num_owned_nodes, int *num_local_nodes) {
  CHECK_HANDLE(f);
  if (need_nodes(f, "tmrx_forest_get_node_conn")) return -1;
  if (conn) *conn = f->f->conn().data();
  if (num_elements) *num_elements = f->f->num_elements();
  if (num_owned_nodes) *num_owned_nodes = f->f->num_owned_nodes();
  if (num_local_nodes) *num_local_nodes = (int)f->f->node_numbers().size();
  return 0;
}

int tmrx_forest_set_mesh_order(tmrx_forest *f, int mesh_order, int interp_type) {
  CHECK_HANDLE(f);
  if (f->f->set_mesh_order(mesh_order, interp_type)) {
    set_error(f->f->last_error);
    return -1;
  }
  return 0;
}

int tmrx_forest_get_dep_node_conn(tmrx_forest *f, const int **ptr, const int **conn,
                                  const double **weights) {
  CHECK_HANDLE(f);
  if (need_nodes(f, "tmrx_forest_get_dep_node_conn")) return -1;
  if (ptr) *ptr = f->f->dep_ptr().data();
  if (conn) *conn = f->f->dep_conn().data();
  if (weights) *weights = f->f->dep_weights().data();
  return f->f->num_dep_nodes();
}

int tmrx_forest_get_owned_node_range(tmrx_forest *f, int range[2]) {
  CHECK_HANDLE(f);
  if (need_nodes(f, "tmrx_forest_get_owned_node_range")) return -1;
  range[0] = 0;
  range[1] = f->f->num_owned_nodes();
  return 0;
}

int tmrx_forest_get_node_numbers(tmrx_forest *f, const int **nums, int *n) {
  CHECK_HANDLE(f);
  if (need_nodes(f, "tmrx_forest_get_node_numbers")) return -1;
  if (nums) *nums = f->f->node_numbers().data();
  if (n) *n = (int)f->f->node_numbers().size();
  return 0;
}

int tmrx_forest_get_points(tmrx_forest *f, const double **X, int *n) {
  CHECK_HANDLE(f);
  if (need_nodes(f, "tmrx_forest_get_points")) return -1;
  if (!f->f->has_geometry()) {
    set_error("tmrx_forest_get_points: no block points were set");
    return -1;
  }
  if (X) *X = f->f->points().data();
  if (n) *n = (int)f->f->node_numbers().size();
  return 0;
}

int tmrx_forest_get_nodes_with_name(tmrx_forest *f, const char *name, int *out, int cap) {
  CHECK_HANDLE(f);
  if (need_nodes(f, "tmrx_forest_get_nodes_with_name")) return -1;
  std::vector<int> v;
  f->f->nodes_with_name(name, v);
  int n = (int)v.size();
  if (out && cap > 0) std::memcpy(out, v.data(), sizeof(int) * (size_t)(n < cap ? n : cap));
  return n;
}

int tmrx_forest_get_octs_with_name(tmrx_forest *f, const char *name, tmrx_octant *out,
                                   int cap) {
  CHECK_HANDLE(f);
  std::vector<tmrx::Oct> v;
  f->f->octs_with_name(name, v);
  int n = (int)v.size();
  if (out && cap > 0) std::memcpy(out, v.data(), sizeof(tmrx_octant) * (size_t)(n < cap ? n : cap));
  return n;
}

int tmrx_forest_create_interpolation(tmrx_forest *fine, tmrx_forest *coarse, int *nrows,
                                     int *ncols, const int **rowp, const int **cols,
                                     const double **weights) {
  CHECK_HANDLE(fine);
  CHECK_HANDLE(coarse);
  if (fine->f->create_interpolation(coarse->f, fine->P)) return fail_from(fine);
  if (nrows) *nrows = fine->P.nrows;
  if (ncols) *ncols = fine->P.ncols;
  if (rowp) *rowp = fine->P.rowp.data();
  if (cols) *cols = fine->P.cols.data();
  if (weights) *weights = fine->P.w.data();
  return 0;
}

int tmrx_forest_create_interpolation_device(tmrx_forest *fine, tmrx_forest *coarse, int device,
                                            int *nrows, int *ncols, const int **rowp,
                                            const int **cols, const double **weights, double *ms) {
  CHECK_HANDLE(fine);
  CHECK_HANDLE(coarse);
  if (fine->f->create_interpolation_device(coarse->f, fine->P, device, ms)) return fail_from(fine);
  if (nrows) *nrows = fine->P.nrows;
  if (ncols) *ncols = fine->P.ncols;
  if (rowp) *rowp = fine->P.rowp.data();
  if (cols) *cols = fine->P.cols.data();
  if (weights) *weights = fine->P.w.data();
  return 0;
}

int tmrx_forest_create_interpolation_geometric(tmrx_forest *fine, tmrx_forest *coarse, int *nrows,
                                               int *ncols, const int **rowp, const int **cols,
                                               const double **weights) {
  CHECK_HANDLE(fine);
  CHECK_HANDLE(coarse);
  if (fine->f->create_interpolation_geometric(coarse->f, fine->P)) return fail_from(fine);
  if (nrows) *nrows = fine->P.nrows;
  if (ncols) *ncols = fine->P.ncols;
  if (rowp) *rowp = fine->P.rowp.data();
  if (cols) *cols = fine->P.cols.data();
  if (weights) *weights = fine->P.w.data();
  return 0;
}

int tmrx_forest_element_center_weights(tmrx_forest *f, tmrx_forest *filter, int *nrows,
                                       int *ncols, const int **rowp, const int **cols,
                                       const double **weights) {
  CHECK_HANDLE(f);
  CHECK_HANDLE(filter);
  if (f->f->create_nodes() != 0) return fail_from(f);
  if (f->f->element_center_weights(filter->f, f->P)) return fail_from(f);
  if (nrows) *nrows = f->P.nrows;
  if (ncols) *ncols = f->P.ncols;
  if (rowp) *rowp = f->P.rowp.data();
  if (cols) *cols = f->P.cols.data();
  if (weights) *weights = f->P.w.data();
  return 0;
}

}  // extern "C"

// accessor for the solver side
tmrx::Forest *tmrx_forest_impl(tmrx_forest *f) { return f ? f->f : nullptr; }
