// Host-side forest of octrees for the hex8 topology-optimisation path.
//
// One process owns the whole forest (the GPU it feeds holds the full
// problem), so the MPI ownership machinery of TMROctForest collapses to a
// single rank.  Everything that decides an index -- Morton order, 2:1 balance,
// node numbering, dependent-node weights, interpolation rows -- follows the
// semantics of src/TMROctForest.cpp so that the maps handed to the GPU are the
// ones TMR would produce for the same input.
//
// Ordering is computed with 128-bit interleaved keys (block | z-order of the
// 30-bit coordinates with x as the most significant bit of each triple), which
// sorts exactly like TMROctant::compare (src/TMROctant.cpp:171-204).
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace tmrx {

constexpr int MAX_LEVEL = 30;                 // src/TMRBase.h:37
constexpr int32_t HMAX = int32_t(1) << MAX_LEVEL;

// Same field layout as TMROctant (src/TMROctant.h:48-53).
struct Oct {
  int32_t block;
  int32_t x, y, z;
  int32_t tag;
  int16_t level;
  int16_t info;
};
static_assert(sizeof(Oct) == 24, "octant must stay 24 bytes");

using u128 = unsigned __int128;

// Position key: block in the top 32 bits, then the z-order interleave of the
// 32-bit coordinates (valid for 0 <= x, y, z <= 2^30).
u128 pos_key(int32_t block, int32_t x, int32_t y, int32_t z);
inline u128 pos_key(const Oct &o) { return pos_key(o.block, o.x, o.y, o.z); }

int child_id(const Oct &o);
Oct sibling(const Oct &o, int id);
Oct parent_of(const Oct &o);
Oct face_neighbor(const Oct &o, int face);
Oct edge_neighbor(const Oct &o, int edge);
Oct corner_neighbor(const Oct &o, int corner);

// Block (coarse hex) connectivity: TMRBlockConn restated.
struct BlockConn {
  int num_nodes = 0, num_edges = 0, num_faces = 0, num_blocks = 0;
  std::vector<int> block_conn, block_edge_conn, block_face_conn, block_face_ids;
  std::vector<int> node_block_ptr, node_block_conn;
  std::vector<int> edge_block_ptr, edge_block_conn;
  std::vector<int> face_block_ptr, face_block_conn;
  std::vector<int> face_block_owners, edge_block_owners, node_block_owners;
};

// CSR interpolation operator (rows = fine independent nodes).
struct InterpCSR {
  int nrows = 0, ncols = 0;
  std::vector<int> rowp, cols;
  std::vector<double> w;
};

class Forest {
 public:
  Forest(int order = 2, int interp = 1);

  // connectivity / geometry / names ---------------------------------------
  int set_connectivity(int num_nodes, const int *block_conn, int num_blocks);
  int set_node_points(const double *X);  // num_nodes x 3 block vertices
  int set_name(int kind, int index, const char *name);  // 0 vtx,1 edge,2 face,3 block
  const BlockConn &conn_data() const { return bd_; }

  // octree operations ------------------------------------------------------
  // uniform trees of depth `level`, or of depth block_levels[b] in block b
  int create_trees(int level, const int *block_levels = nullptr);
  // octants of a saved forest (validated: aligned, inside their blocks,
  // tiling every block exactly)
  int set_octants(const Oct *o, int n);
  int refine(const int *refinement, int min_level, int max_level);
  int balance(int balance_corner);
  // the same balance on GPU `device`, bit for bit (octants, tags, order)
  int balance_device(int balance_corner, int device, double *ms = nullptr);
  Forest *coarsen() const;
  Forest *duplicate() const;
  int create_nodes();
  // create_nodes on GPU `device`, the same results bit for bit; ms (may be
  // null) receives the wall time of the device pass with its copies
  int create_nodes_device(int device, double *ms = nullptr);
  // TMROctForest::setMeshOrder: orders below 2 become 2 (the reference
  // clamps to [2, MAX_ORDER]); only order 2 builds nodes here, higher orders
  // are an error.  Frees the nodes (createNodes builds them again).
  int set_mesh_order(int order, int interp);

  // queries ------------------------------------------------------------------
  const std::vector<Oct> &octants() const { return octs_; }
  bool has_nodes() const { return nodes_ready_; }
  int num_elements() const { return (int)octs_.size(); }
  int num_owned_nodes() const { return num_owned_; }
  int num_dep_nodes() const { return num_dep_; }
  const std::vector<int> &conn() const { return conn_; }
  const std::vector<int> &dep_ptr() const { return dep_ptr_; }
  const std::vector<int> &dep_conn() const { return dep_conn_; }
  const std::vector<double> &dep_weights() const { return dep_w_; }
  const std::vector<int> &node_numbers() const { return node_numbers_; }
  const std::vector<double> &points() const { return X_; }
  int mesh_order() const { return order_; }
  int interp_type() const { return interp_; }
  bool has_geometry() const { return !Xv_.empty(); }

  int nodes_with_name(const char *name, std::vector<int> &out) const;
  int octs_with_name(const char *name, std::vector<Oct> &out) const;
  int create_interpolation(Forest *coarse, InterpCSR &P);
  // the same on GPU `device` (nodes of both forests built there if missing)
  int create_interpolation_device(Forest *coarse, InterpCSR &P, int device, double *ms = nullptr);
  // index of the octant whose region contains the lower corner of o, or -1
  int locate(const Oct &o) const;
  // element-centre interpolation weights from a filter forest: rows =
  // elements of this forest, columns = independent filter nodes
  int element_center_weights(Forest *filter, InterpCSR &W);

  // physical size of an element along each axis (box blocks only)
  void element_extent(const Oct &o, double h[3]) const;
  bool blocks_are_boxes() const;

  // Coarse spaces that are not nested (box blocks only; nonnested.cpp).
  // True if both forests have the same blocks (connectivity and vertex
  // points), i.e. one can be a coarsen() of the other.
  bool same_blocks(const Forest *o) const;
  // Trilinear interpolation from `coarse` to this forest's independent nodes
  // by physical position: each fine node is located in the coarse forest
  // (lowest-index block containing it, then the octant), and the coarse
  // element's shape functions there give the row, coarse hanging nodes
  // expanded, merged and sorted by column.  The forests may have unrelated
  // blocks over the same region.
  int create_interpolation_geometric(Forest *coarse, InterpCSR &P);
  // index of the coarse element containing each element centre of this
  // forest (-1: outside the coarse forest)
  int locate_centres(Forest *coarse, std::vector<int> &owner);

  std::string last_error;

 private:
  void free_mesh();
  void compute_connectivity();
  void transform_node(Oct *oct) const;
  void add_face_neighbors(int face_index, Oct p, std::vector<Oct> &out) const;
  void add_edge_neighbors(int edge_index, Oct p, std::vector<Oct> &out) const;
  void add_corner_neighbors(int corner, Oct p, std::vector<Oct> &out) const;
  void balance_octant(const Oct &oct, int balance_corner, std::vector<Oct> &out) const;
  bool contains_exact(const Oct &o) const;
  bool check_adjacent_faces(int face_index, const Oct &nb) const;
  bool check_adjacent_edges(int edge_index, const Oct &nb) const;
  void compute_dep_faces_and_edges();
  int find_node(const Oct &n) const;
  const Oct *find_enclosing(const Oct &node) const;
  void eval_points();

  int order_, interp_;
  double knots_[2];
  BlockConn bd_;
  std::vector<double> Xv_;  // block vertex points
  std::vector<std::string> names_[4];

  std::vector<Oct> octs_;   // sorted, tag == index
  std::vector<u128> okey_;  // sorted position keys of octs_ (for lookups)

  bool nodes_ready_ = false;
  std::vector<Oct> nodes_;  // sorted unique node octants
  std::vector<u128> nkey_;
  std::vector<int> conn_, node_numbers_, dep_ptr_, dep_conn_;
  std::vector<double> dep_w_, X_;
  int num_owned_ = 0, num_dep_ = 0;
};

}  // namespace tmrx
