// Host-side octree forest (single process).  See forest.h.
//
// Reference semantics followed (file:line in smdogroup/tmr):
//   octant algebra           src/TMROctant.cpp:28-162
//   Morton order             src/TMROctant.cpp:171-274
//   block connectivity       src/TMROctForest.cpp:558-1143
//   createTrees / refine     src/TMROctForest.cpp:1744-1833, 2169-2329
//   coarsen / duplicate      src/TMROctForest.cpp:2097-2164
//   balance                  src/TMROctForest.cpp:2525-3089
//   createNodes              src/TMROctForest.cpp:3465-4867
//   dependent connectivity   src/TMROctForest.cpp:4886-5519
//   node locations           src/TMROctForest.cpp:5524-5680
//   names                    src/TMROctForest.cpp:5747-6203
//   interpolation            src/TMROctForest.cpp:6228-6793
#include "forest.h"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace tmrx {

// Corner/edge/face numbering conventions of the hex block and octant (these
// define the meaning of every index handed across the boundary, so they are
// the same tables as src/TMROctForest.cpp:28-62).
static const int kEdgeNodes[12][2] = {{0, 1}, {2, 3}, {4, 5}, {6, 7}, {0, 2}, {1, 3},
                                      {4, 6}, {5, 7}, {0, 4}, {1, 5}, {2, 6}, {3, 7}};
static const int kFaceNodes[6][4] = {{0, 2, 4, 6}, {1, 3, 5, 7}, {0, 1, 4, 5},
                                     {2, 3, 6, 7}, {0, 1, 2, 3}, {4, 5, 6, 7}};
static const int kFaceEdges[6][4] = {{8, 10, 4, 6}, {9, 11, 5, 7}, {8, 9, 0, 2},
                                     {10, 11, 1, 3}, {4, 5, 0, 1},  {6, 7, 2, 3}};
static const int kChildFaces[8][3] = {{0, 2, 4}, {1, 2, 4}, {0, 3, 4}, {1, 3, 4},
                                      {0, 2, 5}, {1, 2, 5}, {0, 3, 5}, {1, 3, 5}};
static const int kChildFaceEdges[8][3][2] = {
    {{8, 4}, {8, 0}, {4, 0}},   {{9, 5}, {9, 0}, {5, 0}},   {{10, 4}, {10, 1}, {4, 1}},
    {{11, 5}, {11, 1}, {5, 1}}, {{8, 6}, {8, 2}, {6, 2}},   {{9, 7}, {9, 2}, {7, 2}},
    {{10, 6}, {10, 3}, {6, 3}}, {{11, 7}, {11, 3}, {7, 3}}};
static const int kChildEdges[8][3] = {{0, 4, 8},  {0, 5, 9},  {1, 4, 10}, {1, 5, 11},
                                      {2, 6, 8},  {2, 7, 9},  {3, 6, 10}, {3, 7, 11}};
static const int kFaceOrient[8][4] = {{0, 1, 2, 3}, {2, 0, 3, 1}, {3, 2, 1, 0},
                                      {1, 3, 0, 2}, {0, 2, 1, 3}, {2, 3, 0, 1},
                                      {3, 1, 2, 0}, {1, 0, 3, 2}};

// ---------------------------------------------------------------------------
// keys
static inline uint64_t spread16(uint32_t v) {
  uint64_t x = v & 0xffffu;
  x = (x | (x << 32)) & 0x1f00000000ffffULL;
  x = (x | (x << 16)) & 0x1f0000ff0000ffULL;
  x = (x | (x << 8)) & 0x100f00f00f00f00fULL;
  x = (x | (x << 4)) & 0x10c30c30c30c30c3ULL;
  x = (x | (x << 2)) & 0x1249249249249249ULL;
  return x;
}

u128 pos_key(int32_t block, int32_t x, int32_t y, int32_t z) {
  uint32_t ux = (uint32_t)x, uy = (uint32_t)y, uz = (uint32_t)z;
  uint64_t lo = (spread16(ux) << 2) | (spread16(uy) << 1) | spread16(uz);
  uint64_t hi = (spread16(ux >> 16) << 2) | (spread16(uy >> 16) << 1) | spread16(uz >> 16);
  return ((u128)(uint32_t)block << 96) | ((u128)hi << 48) | (u128)lo;
}

// ---------------------------------------------------------------------------
// octant algebra (src/TMROctant.cpp:28-162)
int child_id(const Oct &o) {
  const int32_t h = 1 << (MAX_LEVEL - o.level);
  return ((o.x & h) ? 1 : 0) | ((o.y & h) ? 2 : 0) | ((o.z & h) ? 4 : 0);
}

Oct sibling(const Oct &o, int id) {
  const int32_t h = 1 << (MAX_LEVEL - o.level);
  int32_t xr = (o.x & h) ? o.x - h : o.x;
  int32_t yr = (o.y & h) ? o.y - h : o.y;
  int32_t zr = (o.z & h) ? o.z - h : o.z;
  Oct s = o;
  s.info = 0;
  s.x = (id & 1) ? xr + h : xr;
  s.y = (id & 2) ? yr + h : yr;
  s.z = (id & 4) ? zr + h : zr;
  return s;
}

Oct parent_of(const Oct &o) {
  Oct p = o;
  p.info = 0;
  if (o.level > 0) {
    const int32_t h = 1 << (MAX_LEVEL - o.level);
    p.level = o.level - 1;
    p.x = o.x & ~h;
    p.y = o.y & ~h;
    p.z = o.z & ~h;
  }
  return p;
}

Oct face_neighbor(const Oct &o, int face) {
  const int32_t h = 1 << (MAX_LEVEL - o.level);
  Oct n = o;
  n.info = 0;
  n.x = o.x + ((face == 0) ? -h : (face == 1) ? h : 0);
  n.y = o.y + ((face == 2) ? -h : (face == 3) ? h : 0);
  n.z = o.z + ((face == 4) ? -h : (face == 5) ? h : 0);
  return n;
}

Oct edge_neighbor(const Oct &o, int edge) {
  const int32_t h = 1 << (MAX_LEVEL - o.level);
  Oct n = o;
  n.info = 0;
  if (edge < 4) {
    n.y = (edge % 2 == 0) ? o.y - h : o.y + h;
    n.z = (edge < 2) ? o.z - h : o.z + h;
  } else if (edge < 8) {
    n.x = (edge % 2 == 0) ? o.x - h : o.x + h;
    n.z = (edge < 6) ? o.z - h : o.z + h;
  } else {
    n.x = (edge % 2 == 0) ? o.x - h : o.x + h;
    n.y = (edge < 10) ? o.y - h : o.y + h;
  }
  return n;
}

Oct corner_neighbor(const Oct &o, int corner) {
  const int32_t h = 1 << (MAX_LEVEL - o.level);
  Oct n = o;
  n.info = 0;
  n.x = o.x + (2 * (corner & 1) - 1) * h;
  n.y = o.y + ((corner & 2) - 1) * h;
  n.z = o.z + ((corner & 4) / 2 - 1) * h;
  return n;
}

static inline bool oct_less(const Oct &a, const Oct &b) {
  u128 ka = pos_key(a), kb = pos_key(b);
  if (ka != kb) return ka < kb;
  return a.level < b.level;
}

// Sort by position then level and keep the finest octant per position
// (TMROctantArray::sort, src/TMROctant.cpp:357-399, use_node_index = 0).
static void sort_unique_finest(std::vector<Oct> &v) {
  struct KL {
    u128 k;
    int16_t level;
    int32_t idx;
  };
  std::vector<KL> kl(v.size());
  for (size_t i = 0; i < v.size(); i++) kl[i] = {pos_key(v[i]), v[i].level, (int32_t)i};
  std::sort(kl.begin(), kl.end(), [](const KL &a, const KL &b) {
    if (a.k != b.k) return a.k < b.k;
    return a.level < b.level;
  });
  std::vector<Oct> out;
  out.reserve(v.size());
  for (size_t i = 0; i < kl.size(); i++) {
    if (i + 1 < kl.size() && kl[i + 1].k == kl[i].k) continue;
    out.push_back(v[kl[i].idx]);
  }
  v.swap(out);
}

// ---------------------------------------------------------------------------
Forest::Forest(int order, int interp) : order_(order), interp_(interp) {
  knots_[0] = -1.0;
  knots_[1] = 1.0;
}

void Forest::free_mesh() {
  nodes_ready_ = false;
  nodes_.clear();
  nkey_.clear();
  conn_.clear();
  node_numbers_.clear();
  dep_ptr_.clear();
  dep_conn_.clear();
  dep_w_.clear();
  X_.clear();
  num_owned_ = num_dep_ = 0;
}

static void reindex(const std::vector<Oct> &o, std::vector<u128> &k) {
  k.resize(o.size());
  for (size_t i = 0; i < o.size(); i++) k[i] = pos_key(o[i]);
}

// ---------------------------------------------------------------------------
// Block connectivity (src/TMROctForest.cpp:558-1143)
int Forest::set_connectivity(int num_nodes, const int *block_conn, int num_blocks) {
  if (num_nodes <= 0 || num_blocks <= 0 || !block_conn) {
    last_error = "set_connectivity: empty connectivity";
    return -1;
  }
  for (int i = 0; i < 8 * num_blocks; i++) {
    if (block_conn[i] < 0 || block_conn[i] >= num_nodes) {
      last_error = "set_connectivity: block vertex index out of range";
      return -1;
    }
  }
  bd_ = BlockConn();
  bd_.num_nodes = num_nodes;
  bd_.num_blocks = num_blocks;
  bd_.block_conn.assign(block_conn, block_conn + 8 * num_blocks);
  for (auto &n : names_) n.clear();
  octs_.clear();
  okey_.clear();
  free_mesh();
  compute_connectivity();
  names_[0].assign(bd_.num_nodes, std::string());
  names_[1].assign(bd_.num_edges, std::string());
  names_[2].assign(bd_.num_faces, std::string());
  names_[3].assign(bd_.num_blocks, std::string());
  return 0;
}

void Forest::compute_connectivity() {
  BlockConn &b = bd_;
  const int nb = b.num_blocks, nn = b.num_nodes;
  // node -> block (computeNodesToBlocks)
  b.node_block_ptr.assign(nn + 1, 0);
  for (int i = 0; i < 8 * nb; i++) b.node_block_ptr[b.block_conn[i] + 1]++;
  for (int i = 0; i < nn; i++) b.node_block_ptr[i + 1] += b.node_block_ptr[i];
  b.node_block_conn.assign(b.node_block_ptr[nn], 0);
  {
    std::vector<int> fill(b.node_block_ptr.begin(), b.node_block_ptr.end() - 1);
    for (int i = 0; i < nb; i++)
      for (int j = 0; j < 8; j++) {
        int node = b.block_conn[8 * i + j];
        int adj_index = 0;
        for (; adj_index < 8; adj_index++)
          if (b.block_conn[8 * i + adj_index] == node) break;
        b.node_block_conn[fill[node]++] = 8 * i + adj_index;
      }
  }
  // unique edge numbering (computeEdgesFromNodes)
  b.block_edge_conn.assign(12 * nb, -1);
  int edge = 0;
  for (int i = 0; i < nb; i++)
    for (int j = 0; j < 12; j++) {
      if (b.block_edge_conn[12 * i + j] >= 0) continue;
      int n1 = b.block_conn[8 * i + kEdgeNodes[j][0]];
      int n2 = b.block_conn[8 * i + kEdgeNodes[j][1]];
      std::vector<int> idx{12 * i + j};
      int num = -1;
      for (int ip = b.node_block_ptr[n1]; ip < b.node_block_ptr[n1 + 1]; ip++) {
        int ii = b.node_block_conn[ip] / 8;
        for (int jj = 0; jj < 12; jj++) {
          int m1 = b.block_conn[8 * ii + kEdgeNodes[jj][0]];
          int m2 = b.block_conn[8 * ii + kEdgeNodes[jj][1]];
          if ((n1 == m1 && n2 == m2) || (n1 == m2 && n2 == m1)) {
            if (b.block_edge_conn[12 * ii + jj] >= 0)
              num = b.block_edge_conn[12 * ii + jj];
            else if (idx.size() < 128)
              idx.push_back(12 * ii + jj);
          }
        }
      }
      if (num < 0) num = edge++;
      for (int e : idx) b.block_edge_conn[e] = num;
    }
  b.num_edges = edge;
  // edge -> block (computeEdgesToBlocks)
  const int ne = b.num_edges;
  b.edge_block_ptr.assign(ne + 1, 0);
  for (int i = 0; i < 12 * nb; i++) b.edge_block_ptr[b.block_edge_conn[i] + 1]++;
  for (int i = 0; i < ne; i++) b.edge_block_ptr[i + 1] += b.edge_block_ptr[i];
  b.edge_block_conn.assign(b.edge_block_ptr[ne], 0);
  {
    std::vector<int> fill(b.edge_block_ptr.begin(), b.edge_block_ptr.end() - 1);
    for (int i = 0; i < nb; i++)
      for (int j = 0; j < 12; j++) b.edge_block_conn[fill[b.block_edge_conn[12 * i + j]]++] = i;
  }
  for (int e = 0; e < ne; e++) {
    int owner = nb, owner_index = 0;
    for (int ip = b.edge_block_ptr[e]; ip < b.edge_block_ptr[e + 1]; ip++) {
      int blk = b.edge_block_conn[ip];
      if (blk < owner) {
        owner = blk;
        owner_index = 0;
        for (int j = 0; j < 12; j++, owner_index++)
          if (b.block_edge_conn[12 * blk + j] == e) break;
      }
    }
    int n1 = b.block_conn[8 * owner + kEdgeNodes[owner_index][0]];
    int n2 = b.block_conn[8 * owner + kEdgeNodes[owner_index][1]];
    for (int ip = b.edge_block_ptr[e]; ip < b.edge_block_ptr[e + 1]; ip++) {
      int blk = b.edge_block_conn[ip];
      for (int ei = 0; ei < 12; ei++) {
        int m1 = b.block_conn[8 * blk + kEdgeNodes[ei][0]];
        int m2 = b.block_conn[8 * blk + kEdgeNodes[ei][1]];
        if ((n1 == m1 && n2 == m2) || (n1 == m2 && n2 == m1)) {
          b.edge_block_conn[ip] = 12 * blk + ei;
          break;
        }
      }
    }
  }
  // unique face numbering (computeFacesFromNodes)
  b.block_face_conn.assign(6 * nb, -1);
  int face = 0;
  for (int i = 0; i < nb; i++)
    for (int j = 0; j < 6; j++) {
      if (b.block_face_conn[6 * i + j] >= 0) continue;
      int fn[4];
      for (int k = 0; k < 4; k++) fn[k] = b.block_conn[8 * i + kFaceNodes[j][k]];
      std::vector<int> idx{6 * i + j};
      int num = -1;
      int node = fn[0];
      for (int ip = b.node_block_ptr[node]; ip < b.node_block_ptr[node + 1]; ip++) {
        int ii = b.node_block_conn[ip] / 8;
        if (ii == i) continue;
        for (int jj = 0; jj < 6; jj++) {
          int an[4];
          for (int k = 0; k < 4; k++) an[k] = b.block_conn[8 * ii + kFaceNodes[jj][k]];
          bool eq = false;
          for (int ort = 0; ort < 8 && !eq; ort++)
            eq = fn[0] == an[kFaceOrient[ort][0]] && fn[1] == an[kFaceOrient[ort][1]] &&
                 fn[2] == an[kFaceOrient[ort][2]] && fn[3] == an[kFaceOrient[ort][3]];
          if (eq) {
            if (b.block_face_conn[6 * ii + jj] >= 0)
              num = b.block_face_conn[6 * ii + jj];
            else if (idx.size() < 128)
              idx.push_back(6 * ii + jj);
          }
        }
      }
      if (num < 0) num = face++;
      for (int f : idx) b.block_face_conn[f] = num;
    }
  b.num_faces = face;
  // face -> block and orientations (computeFacesToBlocks)
  const int nf = b.num_faces;
  b.face_block_ptr.assign(nf + 1, 0);
  for (int i = 0; i < 6 * nb; i++) b.face_block_ptr[b.block_face_conn[i] + 1]++;
  for (int i = 0; i < nf; i++) b.face_block_ptr[i + 1] += b.face_block_ptr[i];
  b.face_block_conn.assign(b.face_block_ptr[nf], 0);
  {
    std::vector<int> fill(b.face_block_ptr.begin(), b.face_block_ptr.end() - 1);
    for (int i = 0; i < nb; i++)
      for (int j = 0; j < 6; j++) b.face_block_conn[fill[b.block_face_conn[6 * i + j]]++] = i;
  }
  b.block_face_ids.assign(6 * nb, 0);
  for (int f = 0; f < nf; f++) {
    int owner = nb, owner_index = 0;
    for (int ip = b.face_block_ptr[f]; ip < b.face_block_ptr[f + 1]; ip++) {
      int blk = b.face_block_conn[ip];
      if (blk < owner) {
        owner = blk;
        owner_index = 0;
        for (int j = 0; j < 6; j++, owner_index++)
          if (b.block_face_conn[6 * blk + j] == f) break;
      }
    }
    int on[4];
    for (int k = 0; k < 4; k++) on[k] = b.block_conn[8 * owner + kFaceNodes[owner_index][k]];
    for (int ip = b.face_block_ptr[f]; ip < b.face_block_ptr[f + 1]; ip++) {
      int blk = b.face_block_conn[ip];
      int fi = 0;
      for (; fi < 6; fi++) {
        if (f != b.block_face_conn[6 * blk + fi]) continue;
        int an[4];
        for (int k = 0; k < 4; k++) an[k] = b.block_conn[8 * blk + kFaceNodes[fi][k]];
        bool eq = false;
        for (int ort = 0; ort < 8; ort++) {
          eq = on[0] == an[kFaceOrient[ort][0]] && on[1] == an[kFaceOrient[ort][1]] &&
               on[2] == an[kFaceOrient[ort][2]] && on[3] == an[kFaceOrient[ort][3]];
          if (eq) {
            b.block_face_ids[6 * blk + fi] = ort;
            break;
          }
        }
        if (eq) break;
      }
      b.face_block_conn[ip] = 6 * blk + fi;
    }
  }
  // owners: lowest connected block (computeBlockOwners)
  b.face_block_owners.assign(nf, nb);
  for (int f = 0; f < nf; f++)
    for (int ip = b.face_block_ptr[f]; ip < b.face_block_ptr[f + 1]; ip++)
      b.face_block_owners[f] = std::min(b.face_block_owners[f], b.face_block_conn[ip] / 6);
  b.edge_block_owners.assign(ne, nb);
  for (int e = 0; e < ne; e++)
    for (int ip = b.edge_block_ptr[e]; ip < b.edge_block_ptr[e + 1]; ip++)
      b.edge_block_owners[e] = std::min(b.edge_block_owners[e], b.edge_block_conn[ip] / 12);
  b.node_block_owners.assign(nn, nb);
  for (int n = 0; n < nn; n++)
    for (int ip = b.node_block_ptr[n]; ip < b.node_block_ptr[n + 1]; ip++)
      b.node_block_owners[n] = std::min(b.node_block_owners[n], b.node_block_conn[ip] / 8);
}

int Forest::set_node_points(const double *X) {
  if (bd_.num_nodes == 0) {
    last_error = "set_node_points: set the connectivity first";
    return -1;
  }
  Xv_.assign(X, X + 3 * bd_.num_nodes);
  if (nodes_ready_) eval_points();
  return 0;
}

int Forest::set_name(int kind, int index, const char *name) {
  if (kind < 0 || kind > 3) {
    last_error = "set_name: kind must be 0 (vertex), 1 (edge), 2 (face) or 3 (block)";
    return -1;
  }
  if (index < 0 || index >= (int)names_[kind].size()) {
    last_error = "set_name: index out of range";
    return -1;
  }
  names_[kind][index] = name ? name : "";
  return 0;
}

// ---------------------------------------------------------------------------
// face coordinate transforms (src/TMROctForest.cpp:67-214)
static inline void get_face_oct_coords(int face_id, int32_t h, int32_t x, int32_t y,
                                       int32_t *u, int32_t *v) {
  const int32_t hm = HMAX;
  switch (face_id) {
    case 0: *u = x; *v = y; break;
    case 1: *u = hm - h - y; *v = x; break;
    case 2: *u = hm - h - x; *v = hm - h - y; break;
    case 3: *u = y; *v = hm - h - x; break;
    case 4: *u = y; *v = x; break;
    case 5: *u = x; *v = hm - h - y; break;
    case 6: *u = hm - h - y; *v = hm - h - x; break;
    default: *u = hm - h - x; *v = y; break;
  }
}
static inline void set_face_oct_coords(int face_id, int32_t h, int32_t u, int32_t v,
                                       int32_t *x, int32_t *y) {
  const int32_t hm = HMAX;
  switch (face_id) {
    case 0: *x = u; *y = v; break;
    case 1: *x = v; *y = hm - h - u; break;
    case 2: *x = hm - h - u; *y = hm - h - v; break;
    case 3: *x = hm - h - v; *y = u; break;
    case 4: *x = v; *y = u; break;
    case 5: *x = u; *y = hm - h - v; break;
    case 6: *x = hm - h - v; *y = hm - h - u; break;
    default: *x = hm - h - u; *y = v; break;
  }
}
static inline void get_face_node_coords(int face_id, int32_t hm, int32_t x, int32_t y,
                                        int32_t *u, int32_t *v) {
  switch (face_id) {
    case 0: *u = x; *v = y; break;
    case 1: *u = hm - y; *v = x; break;
    case 2: *u = hm - x; *v = hm - y; break;
    case 3: *u = y; *v = hm - x; break;
    case 4: *u = y; *v = x; break;
    case 5: *u = x; *v = hm - y; break;
    case 6: *u = hm - y; *v = hm - x; break;
    default: *u = hm - x; *v = y; break;
  }
}

// ---------------------------------------------------------------------------
// createTrees (src/TMROctForest.cpp:1744-1833), single process
int Forest::create_trees(int level, const int *block_levels) {
  if (bd_.num_blocks == 0) {
    last_error = "create_trees: set the connectivity first";
    return -1;
  }
  auto clamp_level = [](int l) { return l < 0 ? 0 : (l >= MAX_LEVEL ? MAX_LEVEL - 1 : l); };
  free_mesh();
  size_t total = 0;
  for (int b = 0; b < bd_.num_blocks; b++) {
    const int lb = clamp_level(block_levels ? block_levels[b] : level);
    total += (size_t)1 << (3 * lb);
  }
  octs_.resize(total);
  // Morton index m of a block's octant: bit 3k + 2 of m is bit k of x / h,
  // bit 3k + 1 of y, bit 3k of z (the key's x > y > z priority), so the
  // octants come out already in key order
  size_t i = 0;
  for (int b = 0; b < bd_.num_blocks; b++) {
    const int lb = clamp_level(block_levels ? block_levels[b] : level);
    const int32_t h = 1 << (MAX_LEVEL - lb);
    const int64_t nb = int64_t(1) << (3 * lb);
    for (int64_t m = 0; m < nb; m++, i++) {
      int32_t c[3] = {0, 0, 0};
      for (int k = 0; k < lb; k++)
        for (int a = 0; a < 3; a++) c[a] |= (int32_t)((m >> (3 * k + 2 - a)) & 1) << k;
      octs_[i] = {b, c[0] * h, c[1] * h, c[2] * h, (int32_t)i, (int16_t)lb, 0};
    }
  }
  reindex(octs_, okey_);
  return 0;
}

// Octants given by the caller (a forest saved earlier): each inside its
// block and aligned to its level; sorted into Morton order, the octants of
// every block must tile it exactly, each starting where the previous one
// ends on the interleaved index (a level-l octant spans 8^(30 - l) of it).
int Forest::set_octants(const Oct *o, int n) {
  if (bd_.num_blocks == 0) {
    last_error = "set_octants: set the connectivity first";
    return -1;
  }
  if (n <= 0 || !o) {
    last_error = "set_octants: no octants";
    return -1;
  }
  std::vector<Oct> v(o, o + n);
  for (auto &a : v) {
    if (a.block < 0 || a.block >= bd_.num_blocks || a.level < 0 || a.level > MAX_LEVEL) {
      last_error = "set_octants: octant with a block or level out of range";
      return -1;
    }
    const int32_t h = (int32_t)1 << (MAX_LEVEL - a.level);
    const int32_t c[3] = {a.x, a.y, a.z};
    for (int d = 0; d < 3; d++)
      if (c[d] < 0 || c[d] > HMAX - h || c[d] % h != 0) {
        last_error = "set_octants: octant outside its block or not aligned to its level";
        return -1;
      }
    a.info = 0;
  }
  std::sort(v.begin(), v.end(), oct_less);
  const u128 low = (((u128)1) << 96) - 1, full = ((u128)1) << (3 * MAX_LEVEL);
  u128 expect = 0;
  int cur = -1;
  for (const Oct &a : v) {
    if (a.block != cur) {
      if ((cur >= 0 && expect != full) || a.block != cur + 1) {
        last_error = "set_octants: block " + std::to_string(cur + 1) + " is not covered";
        return -1;
      }
      cur = a.block;
      expect = 0;
    }
    if ((pos_key(a) & low) != expect) {
      last_error = "set_octants: octants of block " + std::to_string(cur) +
                   " overlap or leave a gap";
      return -1;
    }
    expect += ((u128)1) << (3 * (MAX_LEVEL - a.level));
  }
  if (cur != bd_.num_blocks - 1 || expect != full) {
    last_error = "set_octants: block " + std::to_string(cur) + " is not covered";
    return -1;
  }
  free_mesh();
  octs_.swap(v);
  for (size_t i = 0; i < octs_.size(); i++) octs_[i].tag = (int32_t)i;
  reindex(octs_, okey_);
  return 0;
}

// refine (src/TMROctForest.cpp:2169-2329), single process
int Forest::refine(const int *refinement, int min_level, int max_level) {
  if (octs_.empty()) {
    last_error = "refine: no octants";
    return -1;
  }
  free_mesh();
  if (min_level < 0) min_level = 0;
  if (max_level > MAX_LEVEL) max_level = MAX_LEVEL;
  if (min_level > max_level) min_level = max_level;
  std::vector<Oct> out;
  out.reserve(octs_.size() * 2);
  for (size_t i = 0; i < octs_.size(); i++) {
    const Oct &a = octs_[i];
    int r = refinement ? refinement[i] : 1;
    if (!refinement) {
      if (a.level < max_level) {
        Oct o = a;
        o.level += 1;
        o.info = 0;
        out.push_back(sibling(o, 0));
      } else {
        out.push_back(a);
      }
      continue;
    }
    if (r == 0) {
      out.push_back(a);
    } else if (r < 0) {
      if (a.level > min_level) {
        int nl = a.level + r;
        if (nl < min_level) nl = min_level;
        Oct o = a;
        o.level = (int16_t)nl;
        o.info = 0;
        const int32_t h = 1 << (MAX_LEVEL - nl);
        o.x -= o.x % h;
        o.y -= o.y % h;
        o.z -= o.z % h;
        out.push_back(o);
      } else {
        out.push_back(a);
      }
    } else {
      if (a.level < max_level) {
        int nl = a.level + r;
        if (nl > max_level) nl = max_level;
        int ref = nl - a.level;
        ref = (ref <= 0) ? 1 : (1 << (ref - 1));
        Oct o = a;
        o.level = (int16_t)nl;
        o.info = 0;
        const int32_t h = 1 << (MAX_LEVEL - nl);
        int32_t x = o.x - (o.x % h), y = o.y - (o.y % h), z = o.z - (o.z % h);
        for (int ii = 0; ii < ref; ii++)
          for (int jj = 0; jj < ref; jj++)
            for (int kk = 0; kk < ref; kk++) {
              o.x = x + 2 * ii * h;
              o.y = y + 2 * jj * h;
              o.z = z + 2 * kk * h;
              out.push_back(o);
            }
      } else {
        out.push_back(a);
      }
    }
  }
  sort_unique_finest(out);
  octs_.swap(out);
  for (size_t i = 0; i < octs_.size(); i++) octs_[i].tag = (int32_t)i;
  reindex(octs_, okey_);
  return 0;
}

Forest *Forest::duplicate() const {
  Forest *f = new Forest(order_, interp_);
  f->bd_ = bd_;
  f->Xv_ = Xv_;
  for (int k = 0; k < 4; k++) f->names_[k] = names_[k];
  f->octs_ = octs_;
  f->okey_ = okey_;
  return f;
}

// coarsen (src/TMROctForest.cpp:2119-2164): the parent of every 0-child, and
// level-0 octants as they are.  The result may overlap; balance() repairs it.
Forest *Forest::coarsen() const {
  Forest *f = new Forest(order_, interp_);
  f->bd_ = bd_;
  f->Xv_ = Xv_;
  for (int k = 0; k < 4; k++) f->names_[k] = names_[k];
  for (const Oct &a : octs_) {
    if (a.level > 0) {
      if (child_id(a) == 0) f->octs_.push_back(parent_of(a));
    } else {
      f->octs_.push_back(a);
    }
  }
  for (size_t i = 0; i < f->octs_.size(); i++) f->octs_[i].tag = (int32_t)i;
  reindex(f->octs_, f->okey_);
  return f;
}

// ---------------------------------------------------------------------------
// balance (src/TMROctForest.cpp:2525-3089)
void Forest::add_face_neighbors(int face_index, Oct p, std::vector<Oct> &out) const {
  const BlockConn &b = bd_;
  int block = p.block;
  int face = b.block_face_conn[6 * block + face_index];
  int face_id = b.block_face_ids[6 * block + face_index];
  const int32_t h = 1 << (MAX_LEVEL - p.level);
  int32_t u, v;
  if (face_index < 2)
    get_face_oct_coords(face_id, 2 * h, p.y, p.z, &u, &v);
  else if (face_index < 4)
    get_face_oct_coords(face_id, 2 * h, p.x, p.z, &u, &v);
  else
    get_face_oct_coords(face_id, 2 * h, p.x, p.y, &u, &v);
  for (int ip = b.face_block_ptr[face]; ip < b.face_block_ptr[face + 1]; ip++) {
    int adj = b.face_block_conn[ip] / 6;
    if (adj == block) continue;
    int ai = b.face_block_conn[ip] % 6;
    int fid = b.block_face_ids[6 * adj + ai];
    Oct n{adj, 0, 0, 0, 0, p.level, 0};
    if (ai < 2) {
      n.x = (HMAX - 2 * h) * (ai % 2);
      set_face_oct_coords(fid, 2 * h, u, v, &n.y, &n.z);
    } else if (ai < 4) {
      n.y = (HMAX - 2 * h) * (ai % 2);
      set_face_oct_coords(fid, 2 * h, u, v, &n.x, &n.z);
    } else {
      n.z = (HMAX - 2 * h) * (ai % 2);
      set_face_oct_coords(fid, 2 * h, u, v, &n.x, &n.y);
    }
    out.push_back(n);
  }
}

void Forest::add_edge_neighbors(int edge_index, Oct p, std::vector<Oct> &out) const {
  const BlockConn &b = bd_;
  int block = p.block;
  int edge = b.block_edge_conn[12 * block + edge_index];
  const int32_t h = 1 << (MAX_LEVEL - p.level);
  int32_t ucoord = (edge_index < 4) ? p.x : (edge_index < 8) ? p.y : p.z;
  int n1 = b.block_conn[8 * block + kEdgeNodes[edge_index][0]];
  int n2 = b.block_conn[8 * block + kEdgeNodes[edge_index][1]];
  for (int ip = b.edge_block_ptr[edge]; ip < b.edge_block_ptr[edge + 1]; ip++) {
    int adj = b.edge_block_conn[ip] / 12;
    if (adj == block) continue;
    int ai = b.edge_block_conn[ip] % 12;
    int m1 = b.block_conn[8 * adj + kEdgeNodes[ai][0]];
    int m2 = b.block_conn[8 * adj + kEdgeNodes[ai][1]];
    int32_t u = (n1 == m2 && n2 == m1) ? HMAX - 2 * h - ucoord : ucoord;
    Oct n{adj, 0, 0, 0, 0, p.level, 0};
    if (ai < 4) {
      n.x = u;
      n.y = (HMAX - 2 * h) * (ai % 2);
      n.z = (HMAX - 2 * h) * (ai / 2);
    } else if (ai < 8) {
      n.x = (HMAX - 2 * h) * (ai % 2);
      n.y = u;
      n.z = (HMAX - 2 * h) * ((ai - 4) / 2);
    } else {
      n.x = (HMAX - 2 * h) * (ai % 2);
      n.y = (HMAX - 2 * h) * ((ai - 8) / 2);
      n.z = u;
    }
    out.push_back(n);
  }
}

void Forest::add_corner_neighbors(int corner, Oct p, std::vector<Oct> &out) const {
  const BlockConn &b = bd_;
  int block = p.block;
  int node = b.block_conn[8 * block + corner];
  const int32_t h = 1 << (MAX_LEVEL - p.level);
  for (int ip = b.node_block_ptr[node]; ip < b.node_block_ptr[node + 1]; ip++) {
    int adj = b.node_block_conn[ip] / 8;
    if (adj == block) continue;
    int ai = b.node_block_conn[ip] % 8;
    Oct n{adj, 0, 0, 0, 0, p.level, 0};
    n.x = (HMAX - 2 * h) * (ai % 2);
    n.y = (HMAX - 2 * h) * ((ai % 4) / 2);
    n.z = (HMAX - 2 * h) * (ai / 4);
    out.push_back(n);
  }
}

static inline bool in_bounds(const Oct &q) {
  return q.x >= 0 && q.x < HMAX && q.y >= 0 && q.y < HMAX && q.z >= 0 && q.z < HMAX;
}

void Forest::balance_octant(const Oct &oct, int balance_corner, std::vector<Oct> &out) const {
  if (oct.level <= 1) return;
  Oct p = parent_of(oct);
  for (int f = 0; f < 6; f++) {
    Oct q = sibling(face_neighbor(p, f), 0);
    if (in_bounds(q))
      out.push_back(q);
    else
      add_face_neighbors(f, q, out);
  }
  for (int e = 0; e < 12; e++) {
    Oct q = sibling(edge_neighbor(p, e), 0);
    if (in_bounds(q)) {
      out.push_back(q);
      continue;
    }
    int ex = (q.x < 0 || q.x >= HMAX), ey = (q.y < 0 || q.y >= HMAX), ez = (q.z < 0 || q.z >= HMAX);
    if ((ex && ey) || (ex && ez) || (ey && ez)) {
      add_edge_neighbors(e, q, out);
    } else {
      int face = ex ? (q.x < 0 ? 0 : 1) : ey ? (q.y < 0 ? 2 : 3) : (q.z < 0 ? 4 : 5);
      add_face_neighbors(face, q, out);
    }
  }
  if (!balance_corner) return;
  for (int c = 0; c < 8; c++) {
    Oct q = sibling(corner_neighbor(p, c), 0);
    if (in_bounds(q)) {
      out.push_back(q);
      continue;
    }
    int ex = (q.x < 0 || q.x >= HMAX), ey = (q.y < 0 || q.y >= HMAX), ez = (q.z < 0 || q.z >= HMAX);
    if (ex && ey && ez) {
      add_corner_neighbors(c, q, out);
    } else if ((ex && ey) || (ex && ez) || (ey && ez)) {
      int edge;
      if (ey && ez)
        edge = (q.y < 0 ? 0 : 1) + (q.z < 0 ? 0 : 2);
      else if (ex && ez)
        edge = (q.x < 0 ? 4 : 5) + (q.z < 0 ? 0 : 2);
      else
        edge = (q.x < 0 ? 8 : 9) + (q.y < 0 ? 0 : 2);
      add_edge_neighbors(edge, q, out);
    } else {
      int face = ex ? (q.x < 0 ? 0 : 1) : ey ? (q.y < 0 ? 2 : 3) : (q.z < 0 ? 4 : 5);
      add_face_neighbors(face, q, out);
    }
  }
}

// The reference closes a hash of 0-children under balanceOctant with a FIFO
// queue.  Every octant balanceOctant adds is one level coarser than the one it
// processes, so the same closure is reached by sweeping the levels from finest
// to coarsest with sorted, de-duplicated per-level sets (no hash, no queue).
int Forest::balance(int balance_corner) {
  if (octs_.empty()) {
    last_error = "balance: no octants have been created";
    return -1;
  }
  free_mesh();
  std::vector<std::vector<Oct>> lev(MAX_LEVEL + 1);
  for (const Oct &a : octs_) {
    Oct s = sibling(a, 0);
    s.tag = 0;
    s.info = 0;
    lev[s.level].push_back(s);
  }
  auto uniq = [](std::vector<Oct> &v) {
    std::vector<std::pair<u128, int>> k(v.size());
    for (size_t i = 0; i < v.size(); i++) k[i] = {pos_key(v[i]), (int)i};
    std::sort(k.begin(), k.end(),
              [](const std::pair<u128, int> &a, const std::pair<u128, int> &b) { return a.first < b.first; });
    std::vector<Oct> o;
    o.reserve(v.size());
    for (size_t i = 0; i < k.size(); i++) {
      if (i > 0 && k[i].first == k[i - 1].first) continue;
      o.push_back(v[k[i].second]);
    }
    v.swap(o);
  };
  for (int L = MAX_LEVEL; L >= 0; L--) {
    if (lev[L].empty()) continue;
    uniq(lev[L]);
    if (L > 1) {
      std::vector<Oct> add;
      add.reserve(lev[L].size() * (balance_corner ? 26 : 18));
      for (const Oct &o : lev[L]) balance_octant(o, balance_corner, add);
      for (Oct &q : add) {
        q.tag = 0;
        q.info = 0;
        lev[L - 1].push_back(q);
      }
    }
  }
  std::vector<Oct> all;
  for (int L = 0; L <= MAX_LEVEL; L++)
    for (const Oct &o : lev[L]) {
      if (o.level > 0) {
        for (int j = 0; j < 8; j++) all.push_back(sibling(o, j));
      } else {
        all.push_back(o);
      }
    }
  sort_unique_finest(all);
  octs_.swap(all);
  for (size_t i = 0; i < octs_.size(); i++) {
    octs_[i].tag = (int32_t)i;
    octs_[i].info = 0;
  }
  reindex(octs_, okey_);
  return 0;
}

// ---------------------------------------------------------------------------
// createNodes and helpers (src/TMROctForest.cpp:3465-4867)
bool Forest::contains_exact(const Oct &o) const {
  u128 k = pos_key(o);
  auto it = std::lower_bound(okey_.begin(), okey_.end(), k);
  if (it == okey_.end() || *it != k) return false;
  return octs_[it - okey_.begin()].level == o.level;
}

bool Forest::check_adjacent_faces(int face_index, const Oct &nb) const {
  const BlockConn &b = bd_;
  const int32_t h = 1 << (MAX_LEVEL - nb.level);
  int owner = nb.block;
  int face = b.block_face_conn[6 * owner + face_index];
  int face_id = b.block_face_ids[6 * owner + face_index];
  int32_t u, v;
  if (face_index < 2)
    get_face_oct_coords(face_id, h, nb.y, nb.z, &u, &v);
  else if (face_index < 4)
    get_face_oct_coords(face_id, h, nb.x, nb.z, &u, &v);
  else
    get_face_oct_coords(face_id, h, nb.x, nb.y, &u, &v);
  for (int ip = b.face_block_ptr[face]; ip < b.face_block_ptr[face + 1]; ip++) {
    int blk = b.face_block_conn[ip] / 6;
    if (blk == owner) continue;
    int ai = b.face_block_conn[ip] % 6;
    int fid = b.block_face_ids[6 * blk + ai];
    Oct o{blk, 0, 0, 0, 0, nb.level, 0};
    if (ai < 2) {
      o.x = (HMAX - h) * (ai % 2);
      set_face_oct_coords(fid, h, u, v, &o.y, &o.z);
    } else if (ai < 4) {
      o.y = (HMAX - h) * (ai % 2);
      set_face_oct_coords(fid, h, u, v, &o.x, &o.z);
    } else {
      o.z = (HMAX - h) * (ai % 2);
      set_face_oct_coords(fid, h, u, v, &o.x, &o.y);
    }
    if (contains_exact(o)) return true;
  }
  return false;
}

bool Forest::check_adjacent_edges(int edge_index, const Oct &nb) const {
  const BlockConn &b = bd_;
  const int32_t h = 1 << (MAX_LEVEL - nb.level);
  int32_t ucoord = (edge_index < 4) ? nb.x : (edge_index < 8) ? nb.y : nb.z;
  int owner = nb.block;
  int edge = b.block_edge_conn[12 * owner + edge_index];
  int n1 = b.block_conn[8 * owner + kEdgeNodes[edge_index][0]];
  int n2 = b.block_conn[8 * owner + kEdgeNodes[edge_index][1]];
  for (int ip = b.edge_block_ptr[edge]; ip < b.edge_block_ptr[edge + 1]; ip++) {
    int blk = b.edge_block_conn[ip] / 12;
    if (blk == owner) continue;
    int ai = b.edge_block_conn[ip] % 12;
    int m1 = b.block_conn[8 * blk + kEdgeNodes[ai][0]];
    int m2 = b.block_conn[8 * blk + kEdgeNodes[ai][1]];
    int32_t u = (n1 == m2 && n2 == m1) ? HMAX - h - ucoord : ucoord;
    Oct o{blk, 0, 0, 0, 0, nb.level, 0};
    if (ai < 4) {
      o.x = u;
      o.y = (HMAX - h) * (ai % 2);
      o.z = (HMAX - h) * (ai / 2);
    } else if (ai < 8) {
      o.x = (HMAX - h) * (ai % 2);
      o.y = u;
      o.z = (HMAX - h) * ((ai - 4) / 2);
    } else {
      o.x = (HMAX - h) * (ai % 2);
      o.y = (HMAX - h) * ((ai - 8) / 2);
      o.z = u;
    }
    if (contains_exact(o)) return true;
  }
  return false;
}

// info bit codec (src/TMROctForest.cpp:218-276)
static inline int encode_info(const Oct &o, int face_info, int edge_info) {
  int id = child_id(o), info = 0;
  for (int k = 0; k < 3; k++) {
    if (face_info & (1 << kChildFaces[id][k])) info |= 1 << k;
    if (edge_info & (1 << kChildEdges[id][k])) info |= 1 << (k + 3);
  }
  return info;
}
static inline void decode_info(const Oct &o, int info, int *face_info, int *edge_info) {
  int id = child_id(o);
  int fi = 0, ei = 0;
  for (int k = 0; k < 3; k++) {
    if (info & (1 << k)) fi |= 1 << kChildFaces[id][k];
    if (info & (1 << (k + 3))) ei |= 1 << kChildEdges[id][k];
  }
  for (int k = 0; k < 3; k++)
    if (info & (1 << k))
      for (int j = 0; j < 2; j++) ei |= 1 << kChildFaceEdges[id][k][j];
  if (face_info) *face_info = fi;
  if (edge_info) *edge_info = ei;
}

void Forest::compute_dep_faces_and_edges() {
  for (Oct &o : octs_) {
    int face_info = 0, edge_info = 0;
    if (o.level == 0) {
      o.info = 0;
      continue;
    }
    int id = child_id(o);
    Oct par = parent_of(o);
    for (int k = 0; k < 3; k++) {
      int fi = kChildFaces[id][k];
      Oct nb = face_neighbor(par, fi);
      if (in_bounds(nb)) {
        if (contains_exact(nb)) face_info |= 1 << fi;
      } else if (check_adjacent_faces(fi, nb)) {
        face_info |= 1 << fi;
      }
    }
    for (int k = 0; k < 3; k++) {
      int ei = kChildEdges[id][k];
      Oct nb = edge_neighbor(par, ei);
      int fx0 = nb.x < 0, fy0 = nb.y < 0, fz0 = nb.z < 0;
      int fx = fx0 || nb.x >= HMAX, fy = fy0 || nb.y >= HMAX, fz = fz0 || nb.z >= HMAX;
      if ((fx && fy) || (fy && fz) || (fx && fz)) {
        if (check_adjacent_edges(ei, nb)) edge_info |= 1 << ei;
      } else if (fx || fy || fz) {
        int fi = fx * (fx0 ? 0 : 1) + fy * (fy0 ? 2 : 3) + fz * (fz0 ? 4 : 5);
        if (check_adjacent_faces(fi, nb)) edge_info |= 1 << ei;
      } else if (contains_exact(nb)) {
        edge_info |= 1 << ei;
      }
    }
    o.info = (int16_t)encode_info(o, face_info, edge_info);
  }
}

// transformNode (src/TMROctForest.cpp:3847-4062) without the edge/face
// orientation outputs, which only higher-order meshes use.
void Forest::transform_node(Oct *oct) const {
  const BlockConn &b = bd_;
  int fx0 = oct->x == 0, fy0 = oct->y == 0, fz0 = oct->z == 0;
  int fx = fx0 || oct->x == HMAX, fy = fy0 || oct->y == HMAX, fz = fz0 || oct->z == HMAX;
  if (!(fx || fy || fz)) return;
  int block = oct->block;
  if (fx && fy && fz) {
    int corner = (fx0 ? 0 : 1) + (fy0 ? 0 : 2) + (fz0 ? 0 : 4);
    int node = b.block_conn[8 * block + corner];
    if (block != b.node_block_owners[node]) {
      int ptr = b.node_block_ptr[node];
      int adj = b.node_block_conn[ptr] / 8, ai = b.node_block_conn[ptr] % 8;
      oct->block = adj;
      oct->x = HMAX * (ai % 2);
      oct->y = HMAX * ((ai % 4) / 2);
      oct->z = HMAX * (ai / 4);
    }
  } else if ((fy && fz) || (fx && fz) || (fx && fy)) {
    int ei;
    int32_t u;
    if (fy && fz) {
      ei = (fy0 ? 0 : 1) + (fz0 ? 0 : 2);
      u = oct->x;
    } else if (fx && fz) {
      ei = (fx0 ? 4 : 5) + (fz0 ? 0 : 2);
      u = oct->y;
    } else {
      ei = (fx0 ? 8 : 9) + (fy0 ? 0 : 2);
      u = oct->z;
    }
    int edge = b.block_edge_conn[12 * block + ei];
    if (block != b.edge_block_owners[edge]) {
      int ptr = b.edge_block_ptr[edge];
      int adj = b.edge_block_conn[ptr] / 12, ai = b.edge_block_conn[ptr] % 12;
      int n1 = b.block_conn[8 * block + kEdgeNodes[ei][0]];
      int n2 = b.block_conn[8 * block + kEdgeNodes[ei][1]];
      int m1 = b.block_conn[8 * adj + kEdgeNodes[ai][0]];
      int m2 = b.block_conn[8 * adj + kEdgeNodes[ai][1]];
      int32_t uo = (n1 == m2 && n2 == m1) ? HMAX - u : u;
      oct->block = adj;
      if (ai < 4) {
        oct->x = uo;
        oct->y = HMAX * (ai % 2);
        oct->z = HMAX * (ai / 2);
      } else if (ai < 8) {
        oct->x = HMAX * (ai % 2);
        oct->y = uo;
        oct->z = HMAX * ((ai - 4) / 2);
      } else {
        oct->x = HMAX * (ai % 2);
        oct->y = HMAX * ((ai - 8) / 2);
        oct->z = uo;
      }
    }
  } else {
    int fi = fx * (fx0 ? 0 : 1) + fy * (fy0 ? 2 : 3) + fz * (fz0 ? 4 : 5);
    int face = b.block_face_conn[6 * block + fi];
    if (block != b.face_block_owners[face]) {
      int face_id = b.block_face_ids[6 * block + fi];
      int32_t u, v;
      if (fi < 2)
        get_face_node_coords(face_id, HMAX, oct->y, oct->z, &u, &v);
      else if (fi < 4)
        get_face_node_coords(face_id, HMAX, oct->x, oct->z, &u, &v);
      else
        get_face_node_coords(face_id, HMAX, oct->x, oct->y, &u, &v);
      int ptr = b.face_block_ptr[face];
      int adj = b.face_block_conn[ptr] / 6, ai = b.face_block_conn[ptr] % 6;
      oct->block = adj;
      if (ai < 2) {
        oct->x = HMAX * (ai % 2);
        oct->y = u;
        oct->z = v;
      } else if (ai < 4) {
        oct->x = u;
        oct->y = HMAX * (ai % 2);
        oct->z = v;
      } else {
        oct->x = u;
        oct->y = v;
        oct->z = HMAX * (ai % 2);
      }
    }
  }
  if (oct->x == HMAX) oct->x = HMAX - 1;
  if (oct->y == HMAX) oct->y = HMAX - 1;
  if (oct->z == HMAX) oct->z = HMAX - 1;
}

int Forest::find_node(const Oct &n) const {
  u128 k = pos_key(n);
  auto it = std::lower_bound(nkey_.begin(), nkey_.end(), k);
  if (it == nkey_.end() || *it != k) return -1;
  return (int)(it - nkey_.begin());
}

// TMROctForest::setMeshOrder (src/TMROctForest.cpp:1389-1427): frees the
// mesh data and sets the order and knots; orders below 2 become 2.  The
// trilinear element is the only one built here.
int Forest::set_mesh_order(int order, int interp) {
  if (order < 2) order = 2;
  if (order != 2) {
    last_error = "set_mesh_order: only mesh order 2 (trilinear hex8) is supported";
    return -1;
  }
  if (interp < 0 || interp > 2) {
    last_error = "set_mesh_order: unknown interpolation type";
    return -1;
  }
  free_mesh();
  order_ = order;
  interp_ = interp;
  return 0;
}

int Forest::create_nodes() {
  if (octs_.empty()) {
    last_error = "create_nodes: no octants have been created";
    return -1;
  }
  if (nodes_ready_) return 0;
  if (order_ != 2) {
    last_error = "create_nodes: only mesh order 2 (trilinear hex8) is supported";
    return -1;
  }
  {
    // the leaves must tile the blocks exactly (refine() leaves holes and
    // coarsen() overlaps until balance() completes the forest)
    u128 vol = 0;
    for (const Oct &o : octs_) vol += (u128)1 << (3 * (MAX_LEVEL - o.level));
    if (vol != (u128)bd_.num_blocks << (3 * MAX_LEVEL)) {
      last_error = "create_nodes: the octants do not tile the blocks; call balance() first";
      return -1;
    }
  }
  compute_dep_faces_and_edges();
  const size_t ne = octs_.size();

  // element corner nodes, canonicalised to the owning block
  struct NK {
    u128 k;
    uint32_t ref;
  };
  std::vector<NK> nk(8 * ne);
  for (size_t i = 0; i < ne; i++) {
    const Oct &o = octs_[i];
    const int32_t h = 1 << (MAX_LEVEL - o.level);
    for (int c = 0; c < 8; c++) {
      Oct n{o.block, o.x + h * (c & 1), o.y + h * ((c >> 1) & 1), o.z + h * (c >> 2), 0, 1, 0};
      transform_node(&n);
      nk[8 * i + c] = {pos_key(n), (uint32_t)(8 * i + c)};
    }
  }
  std::sort(nk.begin(), nk.end(), [](const NK &a, const NK &b) { return a.k < b.k; });
  nodes_.clear();
  nkey_.clear();
  conn_.assign(8 * ne, 0);
  for (size_t i = 0; i < nk.size(); i++) {
    if (i == 0 || nk[i].k != nk[i - 1].k) {
      nkey_.push_back(nk[i].k);
      uint32_t r = nk[i].ref;
      const Oct &o = octs_[r / 8];
      const int32_t h = 1 << (MAX_LEVEL - o.level);
      int c = r % 8;
      Oct n{o.block, o.x + h * (c & 1), o.y + h * ((c >> 1) & 1), o.z + h * (c >> 2), 0, 1, 0};
      transform_node(&n);
      nodes_.push_back(n);
    }
    conn_[nk[i].ref] = (int)nodes_.size() - 1;
  }
  std::vector<NK>().swap(nk);

  // every node a hanging constraint refers to is an element corner of a
  // balanced forest (src/TMROctForest.cpp:4371-4495 adds them otherwise)
  for (size_t i = 0; i < ne; i++) {
    const Oct &o = octs_[i];
    if (!o.info) continue;
    int fi, ei;
    decode_info(o, o.info, &fi, &ei);
    Oct par = parent_of(o);
    const int32_t hp = 1 << (MAX_LEVEL - par.level);
    for (int e = 0; e < 12; e++) {
      if (!(ei & (1 << e))) continue;
      for (int ii = 0; ii < 2; ii++) {
        Oct n{par.block, 0, 0, 0, 0, 1, 0};
        if (e < 4) {
          n.x = par.x + ii * hp; n.y = par.y + hp * (e % 2); n.z = par.z + hp * (e / 2);
        } else if (e < 8) {
          n.x = par.x + hp * (e % 2); n.y = par.y + ii * hp; n.z = par.z + hp * ((e - 4) / 2);
        } else {
          n.x = par.x + hp * (e % 2); n.y = par.y + hp * ((e - 8) / 2); n.z = par.z + ii * hp;
        }
        transform_node(&n);
        if (find_node(n) < 0) {
          last_error = "create_nodes: forest is not 2:1 balanced (missing parent node)";
          free_mesh();
          return -1;
        }
      }
    }
  }

  const int nnodes = (int)nodes_.size();
  std::vector<int> nums(nnodes, 0);
  // labelDependentNodes (src/TMROctForest.cpp:3711-3832), order 2
  for (size_t i = 0; i < ne; i++) {
    const Oct &o = octs_[i];
    if (!o.info) continue;
    int fi, ei;
    decode_info(o, o.info, &fi, &ei);
    for (int f = 0; f < 6; f++)
      if (fi & (1 << f))
        for (int k = 0; k < 4; k++) ei |= 1 << kFaceEdges[f][k];
    const int *c = &conn_[8 * i];
    int id0 = child_id(o);
    for (int e = 0; e < 12; e++) {
      if (!(ei & (1 << e))) continue;
      int id = (e < 4) ? id0 % 2 : (e < 8) ? (id0 % 4) / 2 : id0 / 4;
      int s = (id == 0) ? 1 : 0;  // the node at the middle of the parent edge
      int off;
      if (e < 4)
        off = s + 2 * (e % 2) + 4 * (e / 2);
      else if (e < 8)
        off = (e % 2) + 2 * s + 4 * ((e - 4) / 2);
      else
        off = (e % 2) + 2 * ((e - 8) / 2) + 4 * s;
      nums[c[off]] = -1;
    }
  }
  num_dep_ = 0;
  for (int i = 0; i < nnodes; i++)
    if (nums[i] == -1) nums[i] = -(++num_dep_);

  // createDependentConn (src/TMROctForest.cpp:5157-5519), order 2
  std::vector<int> dptr(num_dep_ + 1, 0);
  for (size_t i = 0; i < ne; i++) {
    const Oct &o = octs_[i];
    if (!o.info) continue;
    int ei;
    decode_info(o, o.info, nullptr, &ei);
    const int *c = &conn_[8 * i];
    for (int e = 0; e < 12; e++) {
      if (!(ei & (1 << e))) continue;
      for (int k = 0; k < 2; k++) {
        int off = (e < 4) ? k + 2 * (e % 2) + 4 * (e / 2)
                          : (e < 8) ? (e % 2) + 2 * k + 4 * ((e - 4) / 2)
                                    : (e % 2) + 2 * ((e - 8) / 2) + 4 * k;
        int idx = nums[c[off]];
        if (idx < 0 && dptr[-idx] == 0) dptr[-idx] = 2;
      }
    }
  }
  for (size_t i = 0; i < ne; i++) {
    const Oct &o = octs_[i];
    if (!o.info) continue;
    int fi;
    decode_info(o, o.info, &fi, nullptr);
    const int *c = &conn_[8 * i];
    for (int f = 0; f < 6; f++) {
      if (!(fi & (1 << f))) continue;
      for (int a = 0; a < 2; a++)
        for (int bb = 0; bb < 2; bb++) {
          int off = (f < 2) ? (f % 2) + 2 * a + 4 * bb
                            : (f < 4) ? a + 2 * (f % 2) + 4 * bb : a + 2 * bb + 4 * (f % 2);
          int idx = nums[c[off]];
          if (idx < 0 && dptr[-idx] == 0) dptr[-idx] = 4;
        }
    }
  }
  for (int i = 0; i < num_dep_; i++) dptr[i + 1] += dptr[i];
  dep_ptr_ = dptr;
  dep_conn_.assign(dptr[num_dep_], 0);
  dep_w_.assign(dptr[num_dep_], 0.0);

  auto lagrange2 = [](double u, double *N) {  // knots {-1, 1}
    N[0] = (u - 1.0) * (1.0 / (-1.0 - 1.0));
    N[1] = (u + 1.0) * (1.0 / (1.0 + 1.0));
  };
  for (size_t i = 0; i < ne; i++) {
    const Oct &o = octs_[i];
    if (!o.info) continue;
    int fi, ei;
    decode_info(o, o.info, &fi, &ei);
    Oct par = parent_of(o);
    const int32_t hp = 1 << (MAX_LEVEL - par.level);
    int id = child_id(o);
    int bx = id % 2, by = (id % 4) / 2, bz = id / 4;
    const int *c = &conn_[8 * i];
    for (int e = 0; e < 12; e++) {
      if (!(ei & (1 << e))) continue;
      int en[2];
      for (int ii = 0; ii < 2; ii++) {  // getEdgeNodes on the parent
        Oct n{par.block, 0, 0, 0, 0, 1, 0};
        if (e < 4) {
          n.x = par.x + ii * hp; n.y = par.y + hp * (e % 2); n.z = par.z + hp * (e / 2);
        } else if (e < 8) {
          n.x = par.x + hp * (e % 2); n.y = par.y + ii * hp; n.z = par.z + hp * ((e - 4) / 2);
        } else {
          n.x = par.x + hp * (e % 2); n.y = par.y + hp * ((e - 8) / 2); n.z = par.z + ii * hp;
        }
        transform_node(&n);
        en[ii] = find_node(n);
      }
      for (int k = 0; k < 2; k++) {
        int off = (e < 4) ? k + 2 * (e % 2) + 4 * (e / 2)
                          : (e < 8) ? (e % 2) + 2 * k + 4 * ((e - 4) / 2)
                                    : (e % 2) + 2 * ((e - 8) / 2) + 4 * k;
        int idx = nums[c[off]];
        if (idx >= 0) continue;
        int d = -idx - 1;
        if (dptr[d + 1] - dptr[d] != 2) continue;
        int ax = (e < 4) ? bx : (e < 8) ? by : bz;
        double u = 1.0 * (ax - 1) + 0.5 * (1.0 + knots_[k]);
        dep_conn_[dptr[d]] = en[0];
        dep_conn_[dptr[d] + 1] = en[1];
        lagrange2(u, &dep_w_[dptr[d]]);
      }
    }
    for (int f = 0; f < 6; f++) {
      if (!(fi & (1 << f))) continue;
      int fn[4];
      for (int jj = 0; jj < 2; jj++)
        for (int ii = 0; ii < 2; ii++) {  // getFaceNodes on the parent
          Oct n{par.block, 0, 0, 0, 0, 1, 0};
          if (f < 2) {
            n.x = par.x + hp * (f % 2); n.y = par.y + hp * ii; n.z = par.z + hp * jj;
          } else if (f < 4) {
            n.x = par.x + hp * ii; n.y = par.y + hp * (f % 2); n.z = par.z + hp * jj;
          } else {
            n.x = par.x + hp * ii; n.y = par.y + hp * jj; n.z = par.z + hp * (f % 2);
          }
          transform_node(&n);
          fn[ii + 2 * jj] = find_node(n);
        }
      for (int jj = 0; jj < 2; jj++)
        for (int ii = 0; ii < 2; ii++) {
          int off = (f < 2) ? (f % 2) + 2 * ii + 4 * jj
                            : (f < 4) ? ii + 2 * (f % 2) + 4 * jj : ii + 2 * jj + 4 * (f % 2);
          int idx = nums[c[off]];
          if (idx >= 0) continue;
          int d = -idx - 1;
          if (dptr[d + 1] - dptr[d] != 4) continue;
          double u = -1.0 + 0.5 * (1.0 + knots_[ii]);
          double v = -1.0 + 0.5 * (1.0 + knots_[jj]);
          if (f < 2) {
            u += 1.0 * by; v += 1.0 * bz;
          } else if (f < 4) {
            u += 1.0 * bx; v += 1.0 * bz;
          } else {
            u += 1.0 * bx; v += 1.0 * by;
          }
          double Nu[2], Nv[2];
          lagrange2(u, Nu);
          lagrange2(v, Nv);
          for (int j = 0; j < 4; j++) {
            dep_conn_[dptr[d] + j] = fn[j];
            dep_w_[dptr[d] + j] = Nu[j % 2] * Nv[j / 2];
          }
        }
    }
  }

  // global numbering: independent nodes in node order, then apply
  num_owned_ = 0;
  for (int i = 0; i < nnodes; i++)
    if (nums[i] >= 0) nums[i] = num_owned_++;
  for (auto &c : conn_) c = nums[c];
  for (auto &c : dep_conn_) c = nums[c];
  node_numbers_ = nums;
  std::sort(node_numbers_.begin(), node_numbers_.end());
  nodes_ready_ = true;
  eval_points();
  return 0;
}

static double convert_to_coordinate(int32_t x) {
  if (x == 0) return 0.0;
  if (x == HMAX - 1) return 1.0;
  return 1.0 * x / HMAX;
}

// evaluateNodeLocations (src/TMROctForest.cpp:5524-5680) with the trilinear
// map of each box block standing in for the topology volume.
void Forest::eval_points() {
  const int nloc = (int)node_numbers_.size();
  X_.assign(3 * (size_t)nloc, 0.0);
  if (Xv_.empty()) return;
  std::vector<char> flag(nloc, 0);
  for (size_t i = 0; i < octs_.size(); i++) {
    const Oct &o = octs_[i];
    const int32_t h = 1 << (MAX_LEVEL - o.level);
    double d = convert_to_coordinate(h);
    double u0 = convert_to_coordinate(o.x), v0 = convert_to_coordinate(o.y),
           w0 = convert_to_coordinate(o.z);
    const int *bc = &bd_.block_conn[8 * o.block];
    for (int c = 0; c < 8; c++) {
      int node = conn_[8 * i + c];
      int index = node >= 0 ? node + num_dep_ : num_dep_ - 1 - (-node - 1);
      if (flag[index]) continue;
      flag[index] = 1;
      double u = u0 + 0.5 * d * (1.0 + knots_[c & 1]);
      double v = v0 + 0.5 * d * (1.0 + knots_[(c >> 1) & 1]);
      double w = w0 + 0.5 * d * (1.0 + knots_[c >> 2]);
      double p[3] = {0, 0, 0};
      for (int k = 0; k < 8; k++) {
        double N = ((k & 1) ? u : 1.0 - u) * ((k & 2) ? v : 1.0 - v) * ((k & 4) ? w : 1.0 - w);
        for (int a = 0; a < 3; a++) p[a] += N * Xv_[3 * bc[k] + a];
      }
      for (int a = 0; a < 3; a++) X_[3 * index + a] = p[a];
    }
  }
}

bool Forest::blocks_are_boxes() const {
  if (Xv_.empty()) return false;
  for (int b = 0; b < bd_.num_blocks; b++) {
    const double *p0 = &Xv_[3 * bd_.block_conn[8 * b]];
    const double *p7 = &Xv_[3 * bd_.block_conn[8 * b + 7]];
    for (int a = 0; a < 3; a++)
      if (!(p7[a] > p0[a])) return false;
    for (int k = 0; k < 8; k++) {
      const double *pk = &Xv_[3 * bd_.block_conn[8 * b + k]];
      for (int a = 0; a < 3; a++) {
        double e = ((k >> a) & 1) ? p7[a] : p0[a];
        if (std::fabs(pk[a] - e) > 1e-12 * (1.0 + std::fabs(e))) return false;
      }
    }
  }
  return true;
}

void Forest::element_extent(const Oct &o, double h[3]) const {
  const double *p0 = &Xv_[3 * bd_.block_conn[8 * o.block]];
  const double *p7 = &Xv_[3 * bd_.block_conn[8 * o.block + 7]];
  double s = std::ldexp(1.0, -o.level);
  for (int a = 0; a < 3; a++) h[a] = (p7[a] - p0[a]) * s;
}

// ---------------------------------------------------------------------------
// getNodesWithName (src/TMROctForest.cpp:5882-6203), order 2.
int Forest::nodes_with_name(const char *name, std::vector<int> &out) const {
  out.clear();
  if (!nodes_ready_) return -1;
  auto match = [&](int kind, int idx) {
    const std::string &s = names_[kind][idx];
    if (!name) return s.empty();
    return !s.empty() && s == name;
  };
  for (size_t i = 0; i < octs_.size(); i++) {
    const Oct &o = octs_[i];
    const int32_t h = 1 << (MAX_LEVEL - o.level);
    int fx0 = o.x == 0, fy0 = o.y == 0, fz0 = o.z == 0;
    int fx1 = o.x + h == HMAX, fy1 = o.y + h == HMAX, fz1 = o.z + h == HMAX;
    int fx = fx0 || fx1, fy = fy0 || fy1, fz = fz0 || fz1;
    const int *c = &conn_[8 * o.tag];
    if (fx && fy && fz) {
      for (int v = 0; v < 8; v++) {
        bool on = ((v & 1) ? fx1 : fx0) && ((v & 2) ? fy1 : fy0) && ((v & 4) ? fz1 : fz0);
        if (on && match(0, bd_.block_conn[8 * o.block + v])) out.push_back(c[v]);
      }
    }
    if ((fy && fz) || (fx && fz) || (fx && fy)) {
      for (int e = 0; e < 12; e++) {
        bool on;
        if (e < 4)
          on = ((e & 1) ? fy1 : fy0) && ((e & 2) ? fz1 : fz0);
        else if (e < 8)
          on = ((e & 1) ? fx1 : fx0) && (((e - 4) & 2) ? fz1 : fz0);
        else
          on = ((e & 1) ? fx1 : fx0) && (((e - 8) & 2) ? fy1 : fy0);
        if (!on || !match(1, bd_.block_edge_conn[12 * o.block + e])) continue;
        for (int k = 0; k < 2; k++) {
          int off = (e < 4) ? k + 2 * (e % 2) + 4 * (e / 2)
                            : (e < 8) ? (e % 2) + 2 * k + 4 * ((e - 4) / 2)
                                      : (e % 2) + 2 * ((e - 8) / 2) + 4 * k;
          out.push_back(c[off]);
        }
      }
    }
    if (fx || fy || fz) {
      const int fon[6] = {fx0, fx1, fy0, fy1, fz0, fz1};
      for (int f = 0; f < 6; f++) {
        if (!fon[f] || !match(2, bd_.block_face_conn[6 * o.block + f])) continue;
        for (int a = 0; a < 2; a++)
          for (int bb = 0; bb < 2; bb++) {
            int off = (f < 2) ? (f % 2) + 2 * bb + 4 * a
                              : (f < 4) ? bb + 2 * (f % 2) + 4 * a : bb + 2 * a + 4 * (f % 2);
            out.push_back(c[off]);
          }
      }
    }
    // An interior octant of a named volume adds nothing in the reference:
    // its loop never advances the count (src/TMROctForest.cpp:6174-6179).
  }
  std::sort(out.begin(), out.end());
  out.erase(std::unique(out.begin(), out.end()), out.end());
  return (int)out.size();
}

// getOctsWithName (src/TMROctForest.cpp:5747-5862)
int Forest::octs_with_name(const char *name, std::vector<Oct> &out) const {
  out.clear();
  auto match = [&](int kind, int idx) {
    const std::string &s = names_[kind][idx];
    if (!name) return s.empty();
    return !s.empty() && s == name;
  };
  for (size_t i = 0; i < octs_.size(); i++) {
    const Oct &o = octs_[i];
    const int32_t h = 1 << (MAX_LEVEL - o.level);
    if (match(3, o.block)) {
      out.push_back(o);
      continue;
    }
    if (o.level == 0) {
      for (int f = 0; f < 6; f++)
        if (match(2, bd_.block_face_conn[6 * o.block + f])) {
          Oct q = o;
          q.info = (int16_t)f;
          out.push_back(q);
        }
      continue;
    }
    int fx0 = o.x == 0, fy0 = o.y == 0, fz0 = o.z == 0;
    int fx = fx0 || o.x + h == HMAX, fy = fy0 || o.y + h == HMAX, fz = fz0 || o.z + h == HMAX;
    const int fl[3] = {fx, fy, fz}, f0[3] = {fx0, fy0, fz0};
    for (int a = 0; a < 3; a++) {
      if (!fl[a]) continue;
      int f = 2 * a + (f0[a] ? 0 : 1);
      if (match(2, bd_.block_face_conn[6 * o.block + f])) {
        Oct q = o;
        q.info = (int16_t)f;
        out.push_back(q);
      }
    }
  }
  return (int)out.size();
}

// ---------------------------------------------------------------------------
// findEnclosing (src/TMROctForest.cpp:6228-6377), order 2, one process
static inline int cmp_pos(const Oct &a, const Oct &b) {
  u128 ka = pos_key(a), kb = pos_key(b);
  return ka < kb ? -1 : (ka > kb ? 1 : 0);
}
static inline bool oct_contains(const Oct &a, const Oct &b) {
  const int32_t h = 1 << (MAX_LEVEL - a.level);
  return b.block == a.block && b.x >= a.x && b.x < a.x + h && b.y >= a.y && b.y < a.y + h &&
         b.z >= a.z && b.z < a.z + h;
}

int Forest::locate(const Oct &o) const {
  u128 k = pos_key(o);
  auto it = std::upper_bound(okey_.begin(), okey_.end(), k);
  if (it == okey_.begin()) return -1;
  int idx = (int)(it - okey_.begin()) - 1;
  return oct_contains(octs_[idx], o) ? idx : -1;
}

const Oct *Forest::find_enclosing(const Oct &node) const {
  const int size = (int)octs_.size();
  const int32_t h = 1 << (MAX_LEVEL - node.level);
  const int ii = node.info % 2, jj = (node.info % 4) / 2, kk = node.info / 4;
  const int32_t xi = node.x + ii * h, yi = node.y + jj * h, zi = node.z + kk * h;
  int low = 0, high = size - 1;
  int mid = low + (high - low) / 2;
  while (mid != low) {
    if (oct_contains(octs_[mid], node)) break;
    int stat = cmp_pos(octs_[mid], node);
    if (stat == 0) break;
    if (stat < 0)
      low = mid + 1;
    else
      high = mid - 1;
    mid = low + (high - low) / 2;
  }
  Oct bound{node.block, node.x + h, node.y + h, node.z + h, 0, 0, 0};
  while (mid < size && cmp_pos(octs_[mid], bound) <= 0) {
    const Oct &a = octs_[mid];
    if (a.block == node.block) {
      const int32_t hm = 1 << (MAX_LEVEL - a.level);
      if (a.x <= xi && xi <= a.x + hm && a.y <= yi && yi <= a.y + hm && a.z <= zi &&
          zi <= a.z + hm)
        return &a;
    }
    mid++;
  }
  return nullptr;
}

// TMROctTACSTopoCreator::createElements/computeWeights
// (src/topology/TMR_TACSTopoCreator.cpp:55-186): the density of an element is
// the filter field at the element centre -- the filter octant containing the
// centre, its trilinear shape functions there, filter hanging nodes expanded,
// merged and sorted by column.  (The reference copies a fixed 8 entries per
// element after the merge; here every merged entry is kept.)
int Forest::element_center_weights(Forest *filter, InterpCSR &W) {
  if (filter->create_nodes() != 0) {
    last_error = "element_center_weights: filter: " + filter->last_error;
    return -1;
  }
  const auto &fo = filter->octs_;
  const auto &fconn = filter->conn_;
  const auto &fdp = filter->dep_ptr_;
  const auto &fdc = filter->dep_conn_;
  const auto &fdw = filter->dep_w_;
  W.nrows = (int)octs_.size();
  W.ncols = filter->num_owned_;
  W.rowp.assign(W.nrows + 1, 0);
  W.cols.clear();
  W.w.clear();
  std::vector<std::pair<int, double>> wts;
  for (size_t i = 0; i < octs_.size(); i++) {
    const Oct &o = octs_[i];
    const int32_t h = 1 << (MAX_LEVEL - o.level);
    Oct c{o.block, o.x + h / 2, o.y + h / 2, o.z + h / 2, 0, (int16_t)MAX_LEVEL, 0};
    if (o.level == MAX_LEVEL) c = o;
    const int t = filter->locate(c);
    if (t < 0) {
      last_error = "element_center_weights: element centre outside the filter forest";
      return -1;
    }
    const Oct &f = fo[t];
    const int32_t hf = 1 << (MAX_LEVEL - f.level);
    const double pc[3] = {-1.0 + 2.0 * ((c.x - f.x) / (double)hf),
                          -1.0 + 2.0 * ((c.y - f.y) / (double)hf),
                          -1.0 + 2.0 * ((c.z - f.z) / (double)hf)};
    wts.clear();
    const int *cc = &fconn[8 * t];
    for (int k = 0; k < 8; k++) {
      double N = 1.0;
      for (int a = 0; a < 3; a++) N *= ((k >> a) & 1) ? 0.5 * (1.0 + pc[a]) : 0.5 * (1.0 - pc[a]);
      const int g = cc[k];
      if (g >= 0) {
        wts.push_back({g, N});
      } else {
        for (int j = fdp[-g - 1]; j < fdp[-g]; j++) wts.push_back({fdc[j], N * fdw[j]});
      }
    }
    std::sort(wts.begin(), wts.end(),
              [](const std::pair<int, double> &a, const std::pair<int, double> &b) { return a.first < b.first; });
    size_t start = W.cols.size();
    for (auto &p : wts) {
      if (W.cols.size() > start && W.cols.back() == p.first)
        W.w.back() += p.second;
      else {
        W.cols.push_back(p.first);
        W.w.push_back(p.second);
      }
    }
    W.rowp[i + 1] = (int)W.cols.size();
  }
  return 0;
}

// createInterpolation + computeElemInterp (src/TMROctForest.cpp:6396-6793)
int Forest::create_interpolation(Forest *coarse, InterpCSR &P) {
  if (coarse == this) {
    last_error = "create_interpolation: cannot interpolate between the same forest";
    return -1;
  }
  if (create_nodes() != 0) return -1;
  if (coarse->create_nodes() != 0) {
    last_error = "create_interpolation: coarse forest: " + coarse->last_error;
    return -1;
  }
  if (coarse->interp_ != interp_) {
    last_error = "create_interpolation: interpolation types differ";
    return -1;
  }
  const int n = num_owned_;
  std::vector<int> rowlen(n, -1);
  std::vector<std::vector<std::pair<int, double>>> rows(n);
  std::vector<std::pair<int, double>> wts;
  const std::vector<int> &cconn = coarse->conn_;
  const std::vector<int> &cptr = coarse->dep_ptr_, &cdc = coarse->dep_conn_;
  const std::vector<double> &cdw = coarse->dep_w_;
  for (size_t i = 0; i < octs_.size(); i++) {
    const int *c = &conn_[8 * i];
    for (int j = 0; j < 8; j++) {
      if (c[j] < 0 || rowlen[c[j]] >= 0) continue;
      Oct node = octs_[i];
      node.info = (int16_t)j;
      const Oct *t = coarse->find_enclosing(node);
      if (!t) {
        last_error = "create_interpolation: no enclosing coarse octant (forests do not nest)";
        return -1;
      }
      const int ii = j % 2, jj = (j % 4) / 2, kk = j / 4;
      const int32_t h = 1 << (MAX_LEVEL - node.level);
      const int32_t hc = 1 << (MAX_LEVEL - t->level);
      double N[3][2];
      int st[3], en[3];
      const int32_t nx[3] = {node.x, node.y, node.z}, tx[3] = {t->x, t->y, t->z};
      const int ijk[3] = {ii, jj, kk};
      for (int a = 0; a < 3; a++) {
        st[a] = 0;
        en[a] = 2;
        if ((ijk[a] == 0 && tx[a] == nx[a]) || (ijk[a] == 1 && tx[a] == nx[a] + h)) {
          st[a] = 0; en[a] = 1; N[a][0] = 1.0;
        } else if ((ijk[a] == 0 && tx[a] + hc == nx[a]) || (ijk[a] == 1 && tx[a] + hc == nx[a] + h)) {
          st[a] = 1; en[a] = 2; N[a][1] = 1.0;
        } else {
          double u = -1.0 + 2.0 * (nx[a] + 0.5 * h * (1.0 + knots_[ijk[a]]) - tx[a]) / hc;
          N[a][0] = (u - 1.0) * (1.0 / (-1.0 - 1.0));
          N[a][1] = (u + 1.0) * (1.0 / (1.0 + 1.0));
        }
      }
      wts.clear();
      const int *cc = &cconn[8 * t->tag];
      for (int z = st[2]; z < en[2]; z++)
        for (int y = st[1]; y < en[1]; y++)
          for (int x = st[0]; x < en[0]; x++) {
            double w = N[0][x] * N[1][y] * N[2][z];
            int cn = cc[x + 2 * y + 4 * z];
            if (cn >= 0) {
              wts.push_back({cn, w});
            } else {
              int d = -cn - 1;
              for (int jp = cptr[d]; jp < cptr[d + 1]; jp++) wts.push_back({cdc[jp], w * cdw[jp]});
            }
          }
      // TMRIndexWeight::uniqueSort (src/TMRBase.h:112-135); stable, so that
      // entries of one column are summed in the order they were formed
      std::stable_sort(wts.begin(), wts.end(),
                [](const std::pair<int, double> &a, const std::pair<int, double> &b) { return a.first < b.first; });
      auto &row = rows[c[j]];
      for (size_t k = 0; k < wts.size(); k++) {
        if (!row.empty() && row.back().first == wts[k].first)
          row.back().second += wts[k].second;
        else
          row.push_back(wts[k]);
      }
      rowlen[c[j]] = (int)row.size();
    }
  }
  P.nrows = n;
  P.ncols = coarse->num_owned_;
  P.rowp.assign(n + 1, 0);
  for (int r = 0; r < n; r++) P.rowp[r + 1] = P.rowp[r] + (int)rows[r].size();
  P.cols.resize(P.rowp[n]);
  P.w.resize(P.rowp[n]);
  for (int r = 0; r < n; r++)
    for (size_t k = 0; k < rows[r].size(); k++) {
      P.cols[P.rowp[r] + k] = rows[r][k].first;
      P.w[P.rowp[r] + k] = rows[r][k].second;
    }
  return 0;
}

}  // namespace tmrx
