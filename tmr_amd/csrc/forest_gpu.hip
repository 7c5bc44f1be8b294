// createNodes on the GPU (TMROctForest::createNodes, src/TMROctForest.cpp:4064-4268,
// with computeDepFacesAndEdges :3619-3702, labelDependentNodes :3711-3832,
// createDependentConn :5157-5519 and the node locations :5524-5680): the same
// octant info bits, node numbering, hanging-node constraints and points as
// Forest::create_nodes, bit for bit, from one thread per octant, corner or
// node and two stable radix passes over the 128-bit node keys.
//
// Where several octants write one entry the host pass keeps a particular one:
// the last octant (in octant order, then edge/face order) for a dependent
// node's constraint, the first octant corner for a node's location.  Here
// every writer posts its rank with an atomic max (min), and only the winner
// writes, so the result does not depend on thread scheduling.
#include <hip/hip_runtime.h>

#include <chrono>
#include <hipcub/hipcub.hpp>
#include <string>
#include <vector>

#include "forest.h"

namespace tmrx {
namespace {

// 128-bit position key as two words: hi = block << 32 | interleaved bits
// 64..95, lo = interleaved bits 0..63 (the host's pos_key)
struct Key {
  uint64_t hi, lo;
};

__device__ __forceinline__ bool key_less(const Key &a, const Key &b) {
  return a.hi < b.hi || (a.hi == b.hi && a.lo < b.lo);
}
__device__ __forceinline__ bool key_eq(const Key &a, const Key &b) {
  return a.hi == b.hi && a.lo == b.lo;
}

__device__ __forceinline__ uint64_t spread16(uint32_t v) {
  uint64_t x = v & 0xffffu;
  x = (x | (x << 32)) & 0x1f00000000ffffULL;
  x = (x | (x << 16)) & 0x1f0000ff0000ffULL;
  x = (x | (x << 8)) & 0x100f00f00f00f00fULL;
  x = (x | (x << 4)) & 0x10c30c30c30c30c3ULL;
  x = (x | (x << 2)) & 0x1249249249249249ULL;
  return x;
}

__device__ __forceinline__ Key make_key(int32_t block, int32_t x, int32_t y, int32_t z) {
  const uint32_t ux = (uint32_t)x, uy = (uint32_t)y, uz = (uint32_t)z;
  const uint64_t lo = (spread16(ux) << 2) | (spread16(uy) << 1) | spread16(uz);
  const uint64_t hi = (spread16(ux >> 16) << 2) | (spread16(uy >> 16) << 1) | spread16(uz >> 16);
  return Key{(hi >> 16) | ((uint64_t)(uint32_t)block << 32), lo | (hi << 48)};
}

// octant or node position (levels as in Oct)
struct P {
  int32_t block, x, y, z, level;
};

__constant__ int cEdgeNodes[12][2] = {{0, 1}, {2, 3}, {4, 5}, {6, 7}, {0, 2}, {1, 3},
                                      {4, 6}, {5, 7}, {0, 4}, {1, 5}, {2, 6}, {3, 7}};
__constant__ int cFaceEdges[6][4] = {{8, 10, 4, 6}, {9, 11, 5, 7}, {8, 9, 0, 2},
                                     {10, 11, 1, 3}, {4, 5, 0, 1},  {6, 7, 2, 3}};
__constant__ int cChildFaces[8][3] = {{0, 2, 4}, {1, 2, 4}, {0, 3, 4}, {1, 3, 4},
                                      {0, 2, 5}, {1, 2, 5}, {0, 3, 5}, {1, 3, 5}};
__constant__ int cChildFaceEdges[8][3][2] = {
    {{8, 4}, {8, 0}, {4, 0}},   {{9, 5}, {9, 0}, {5, 0}},   {{10, 4}, {10, 1}, {4, 1}},
    {{11, 5}, {11, 1}, {5, 1}}, {{8, 6}, {8, 2}, {6, 2}},   {{9, 7}, {9, 2}, {7, 2}},
    {{10, 6}, {10, 3}, {6, 3}}, {{11, 7}, {11, 3}, {7, 3}}};
__constant__ int cChildEdges[8][3] = {{0, 4, 8},  {0, 5, 9},  {1, 4, 10}, {1, 5, 11},
                                      {2, 6, 8},  {2, 7, 9},  {3, 6, 10}, {3, 7, 11}};

// block connectivity on the device (BlockConn)
struct DBlk {
  const int *block_conn, *block_edge_conn, *block_face_conn, *block_face_ids;
  const int *node_block_ptr, *node_block_conn, *edge_block_ptr, *edge_block_conn;
  const int *face_block_ptr, *face_block_conn;
  const int *face_owner, *edge_owner, *node_owner;
};

// sorted octant keys and levels
struct DOcts {
  int n;
  const uint64_t *kh, *kl;
  const int16_t *level;
};

// sorted unique node keys
struct DNodes {
  int n;
  const uint64_t *kh, *kl;
};

__device__ __forceinline__ int lower_bound(int n, const uint64_t *kh, const uint64_t *kl,
                                           const Key &k) {
  int lo = 0, hi = n;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (key_less(Key{kh[mid], kl[mid]}, k))
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

__device__ __forceinline__ bool contains_exact(const DOcts &O, const P &o) {
  const Key k = make_key(o.block, o.x, o.y, o.z);
  const int i = lower_bound(O.n, O.kh, O.kl, k);
  return i < O.n && key_eq(Key{O.kh[i], O.kl[i]}, k) && O.level[i] == o.level;
}

__device__ __forceinline__ int find_node(const DNodes &N, const P &n) {
  const Key k = make_key(n.block, n.x, n.y, n.z);
  const int i = lower_bound(N.n, N.kh, N.kl, k);
  return (i < N.n && key_eq(Key{N.kh[i], N.kl[i]}, k)) ? i : -1;
}

__device__ __forceinline__ int32_t hsize(int level) { return int32_t(1) << (MAX_LEVEL - level); }

__device__ __forceinline__ int child_id(const P &o) {
  const int32_t h = hsize(o.level);
  return ((o.x & h) ? 1 : 0) | ((o.y & h) ? 2 : 0) | ((o.z & h) ? 4 : 0);
}

__device__ __forceinline__ P parent_of(const P &o) {
  P p = o;
  if (o.level > 0) {
    const int32_t h = hsize(o.level);
    p.level = o.level - 1;
    p.x = o.x & ~h;
    p.y = o.y & ~h;
    p.z = o.z & ~h;
  }
  return p;
}

__device__ __forceinline__ P face_neighbor(const P &o, int face) {
  const int32_t h = hsize(o.level);
  P n = o;
  n.x = o.x + ((face == 0) ? -h : (face == 1) ? h : 0);
  n.y = o.y + ((face == 2) ? -h : (face == 3) ? h : 0);
  n.z = o.z + ((face == 4) ? -h : (face == 5) ? h : 0);
  return n;
}

__device__ __forceinline__ P edge_neighbor(const P &o, int edge) {
  const int32_t h = hsize(o.level);
  P n = o;
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

__device__ __forceinline__ bool in_bounds(const P &q) {
  return q.x >= 0 && q.x < HMAX && q.y >= 0 && q.y < HMAX && q.z >= 0 && q.z < HMAX;
}

// face-local coordinates of an octant (size h) or node (hm = HMAX), and back
__device__ __forceinline__ void get_face_coords(int id, int32_t hh, int32_t x, int32_t y,
                                                int32_t &u, int32_t &v) {
  switch (id) {
    case 0: u = x; v = y; break;
    case 1: u = hh - y; v = x; break;
    case 2: u = hh - x; v = hh - y; break;
    case 3: u = y; v = hh - x; break;
    case 4: u = y; v = x; break;
    case 5: u = x; v = hh - y; break;
    case 6: u = hh - y; v = hh - x; break;
    default: u = hh - x; v = y; break;
  }
}
__device__ __forceinline__ void set_face_coords(int id, int32_t hh, int32_t u, int32_t v,
                                                int32_t &x, int32_t &y) {
  switch (id) {
    case 0: x = u; y = v; break;
    case 1: x = v; y = hh - u; break;
    case 2: x = hh - u; y = hh - v; break;
    case 3: x = hh - v; y = u; break;
    case 4: x = v; y = u; break;
    case 5: x = u; y = hh - v; break;
    case 6: x = hh - v; y = hh - u; break;
    default: x = hh - u; y = v; break;
  }
}

// is the octant nb, outside its block across face fi, a leaf of an adjacent block
__device__ bool check_adjacent_faces(const DBlk &B, const DOcts &O, int fi, const P &nb) {
  const int32_t h = hsize(nb.level);
  const int owner = nb.block;
  const int face = B.block_face_conn[6 * owner + fi];
  const int face_id = B.block_face_ids[6 * owner + fi];
  int32_t u, v;
  if (fi < 2)
    get_face_coords(face_id, HMAX - h, nb.y, nb.z, u, v);
  else if (fi < 4)
    get_face_coords(face_id, HMAX - h, nb.x, nb.z, u, v);
  else
    get_face_coords(face_id, HMAX - h, nb.x, nb.y, u, v);
  for (int ip = B.face_block_ptr[face]; ip < B.face_block_ptr[face + 1]; ip++) {
    const int blk = B.face_block_conn[ip] / 6;
    if (blk == owner) continue;
    const int ai = B.face_block_conn[ip] % 6;
    const int fid = B.block_face_ids[6 * blk + ai];
    P o{blk, 0, 0, 0, nb.level};
    if (ai < 2) {
      o.x = (HMAX - h) * (ai % 2);
      set_face_coords(fid, HMAX - h, u, v, o.y, o.z);
    } else if (ai < 4) {
      o.y = (HMAX - h) * (ai % 2);
      set_face_coords(fid, HMAX - h, u, v, o.x, o.z);
    } else {
      o.z = (HMAX - h) * (ai % 2);
      set_face_coords(fid, HMAX - h, u, v, o.x, o.y);
    }
    if (contains_exact(O, o)) return true;
  }
  return false;
}

// the same across block edge ei
__device__ bool check_adjacent_edges(const DBlk &B, const DOcts &O, int ei, const P &nb) {
  const int32_t h = hsize(nb.level);
  const int32_t ucoord = (ei < 4) ? nb.x : (ei < 8) ? nb.y : nb.z;
  const int owner = nb.block;
  const int edge = B.block_edge_conn[12 * owner + ei];
  const int n1 = B.block_conn[8 * owner + cEdgeNodes[ei][0]];
  const int n2 = B.block_conn[8 * owner + cEdgeNodes[ei][1]];
  for (int ip = B.edge_block_ptr[edge]; ip < B.edge_block_ptr[edge + 1]; ip++) {
    const int blk = B.edge_block_conn[ip] / 12;
    if (blk == owner) continue;
    const int ai = B.edge_block_conn[ip] % 12;
    const int m1 = B.block_conn[8 * blk + cEdgeNodes[ai][0]];
    const int m2 = B.block_conn[8 * blk + cEdgeNodes[ai][1]];
    const int32_t u = (n1 == m2 && n2 == m1) ? HMAX - h - ucoord : ucoord;
    P o{blk, 0, 0, 0, nb.level};
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
    if (contains_exact(O, o)) return true;
  }
  return false;
}

// the node's position in its owning block (transformNode, order 2)
__device__ void transform_node(const DBlk &B, P &n) {
  const int fx0 = n.x == 0, fy0 = n.y == 0, fz0 = n.z == 0;
  const int fx = fx0 || n.x == HMAX, fy = fy0 || n.y == HMAX, fz = fz0 || n.z == HMAX;
  if (fx || fy || fz) {
    const int block = n.block;
    if (fx && fy && fz) {
      const int corner = (fx0 ? 0 : 1) + (fy0 ? 0 : 2) + (fz0 ? 0 : 4);
      const int node = B.block_conn[8 * block + corner];
      if (block != B.node_owner[node]) {
        const int ptr = B.node_block_ptr[node];
        const int adj = B.node_block_conn[ptr] / 8, ai = B.node_block_conn[ptr] % 8;
        n.block = adj;
        n.x = HMAX * (ai % 2);
        n.y = HMAX * ((ai % 4) / 2);
        n.z = HMAX * (ai / 4);
      }
    } else if ((fy && fz) || (fx && fz) || (fx && fy)) {
      int ei;
      int32_t u;
      if (fy && fz) {
        ei = (fy0 ? 0 : 1) + (fz0 ? 0 : 2);
        u = n.x;
      } else if (fx && fz) {
        ei = (fx0 ? 4 : 5) + (fz0 ? 0 : 2);
        u = n.y;
      } else {
        ei = (fx0 ? 8 : 9) + (fy0 ? 0 : 2);
        u = n.z;
      }
      const int edge = B.block_edge_conn[12 * block + ei];
      if (block != B.edge_owner[edge]) {
        const int ptr = B.edge_block_ptr[edge];
        const int adj = B.edge_block_conn[ptr] / 12, ai = B.edge_block_conn[ptr] % 12;
        const int n1 = B.block_conn[8 * block + cEdgeNodes[ei][0]];
        const int n2 = B.block_conn[8 * block + cEdgeNodes[ei][1]];
        const int m1 = B.block_conn[8 * adj + cEdgeNodes[ai][0]];
        const int m2 = B.block_conn[8 * adj + cEdgeNodes[ai][1]];
        const int32_t uo = (n1 == m2 && n2 == m1) ? HMAX - u : u;
        n.block = adj;
        if (ai < 4) {
          n.x = uo;
          n.y = HMAX * (ai % 2);
          n.z = HMAX * (ai / 2);
        } else if (ai < 8) {
          n.x = HMAX * (ai % 2);
          n.y = uo;
          n.z = HMAX * ((ai - 4) / 2);
        } else {
          n.x = HMAX * (ai % 2);
          n.y = HMAX * ((ai - 8) / 2);
          n.z = uo;
        }
      }
    } else {
      const int fi = fx * (fx0 ? 0 : 1) + fy * (fy0 ? 2 : 3) + fz * (fz0 ? 4 : 5);
      const int face = B.block_face_conn[6 * block + fi];
      if (block != B.face_owner[face]) {
        const int face_id = B.block_face_ids[6 * block + fi];
        int32_t u, v;
        if (fi < 2)
          get_face_coords(face_id, HMAX, n.y, n.z, u, v);
        else if (fi < 4)
          get_face_coords(face_id, HMAX, n.x, n.z, u, v);
        else
          get_face_coords(face_id, HMAX, n.x, n.y, u, v);
        const int ptr = B.face_block_ptr[face];
        const int adj = B.face_block_conn[ptr] / 6, ai = B.face_block_conn[ptr] % 6;
        n.block = adj;
        if (ai < 2) {
          n.x = HMAX * (ai % 2);
          n.y = u;
          n.z = v;
        } else if (ai < 4) {
          n.x = u;
          n.y = HMAX * (ai % 2);
          n.z = v;
        } else {
          n.x = u;
          n.y = v;
          n.z = HMAX * (ai % 2);
        }
      }
    }
  }
  if (n.x == HMAX) n.x = HMAX - 1;
  if (n.y == HMAX) n.y = HMAX - 1;
  if (n.z == HMAX) n.z = HMAX - 1;
}

__device__ __forceinline__ P load_oct(const Oct *o, int i) {
  const Oct q = o[i];
  return P{q.block, q.x, q.y, q.z, q.level};
}

// corner c of an octant as a node (level 1, as the host's node octants)
__device__ __forceinline__ P corner_node(const P &o, int c) {
  const int32_t h = hsize(o.level);
  return P{o.block, o.x + h * (c & 1), o.y + h * ((c >> 1) & 1), o.z + h * (c >> 2), 1};
}

// node ii of parent edge e (getEdgeNodes) and node (ii, jj) of parent face f
__device__ __forceinline__ P parent_edge_node(const P &par, int e, int ii) {
  const int32_t hp = hsize(par.level);
  P n{par.block, 0, 0, 0, 1};
  if (e < 4) {
    n.x = par.x + ii * hp; n.y = par.y + hp * (e % 2); n.z = par.z + hp * (e / 2);
  } else if (e < 8) {
    n.x = par.x + hp * (e % 2); n.y = par.y + ii * hp; n.z = par.z + hp * ((e - 4) / 2);
  } else {
    n.x = par.x + hp * (e % 2); n.y = par.y + hp * ((e - 8) / 2); n.z = par.z + ii * hp;
  }
  return n;
}
__device__ __forceinline__ P parent_face_node(const P &par, int f, int ii, int jj) {
  const int32_t hp = hsize(par.level);
  P n{par.block, 0, 0, 0, 1};
  if (f < 2) {
    n.x = par.x + hp * (f % 2); n.y = par.y + hp * ii; n.z = par.z + hp * jj;
  } else if (f < 4) {
    n.x = par.x + hp * ii; n.y = par.y + hp * (f % 2); n.z = par.z + hp * jj;
  } else {
    n.x = par.x + hp * ii; n.y = par.y + hp * jj; n.z = par.z + hp * (f % 2);
  }
  return n;
}

// local corner of node k on edge e, node (a, b) on face f
__device__ __forceinline__ int edge_off(int e, int k) {
  return (e < 4) ? k + 2 * (e % 2) + 4 * (e / 2)
                 : (e < 8) ? (e % 2) + 2 * k + 4 * ((e - 4) / 2) : (e % 2) + 2 * ((e - 8) / 2) + 4 * k;
}
__device__ __forceinline__ int face_off(int f, int a, int b) {
  return (f < 2) ? (f % 2) + 2 * a + 4 * b : (f < 4) ? a + 2 * (f % 2) + 4 * b : a + 2 * b + 4 * (f % 2);
}

// info bit codec (face bits 0-2, edge bits 3-5 of the child's faces/edges)
__device__ __forceinline__ void decode_info(int id, int info, int &fi, int &ei) {
  fi = 0;
  ei = 0;
  for (int k = 0; k < 3; k++) {
    if (info & (1 << k)) fi |= 1 << cChildFaces[id][k];
    if (info & (1 << (k + 3))) ei |= 1 << cChildEdges[id][k];
  }
  for (int k = 0; k < 3; k++)
    if (info & (1 << k))
      for (int j = 0; j < 2; j++) ei |= 1 << cChildFaceEdges[id][k][j];
}

// --------------------------------------------------------------------------
__global__ void k_oct_keys(int n, const Oct *__restrict__ octs, uint64_t *kh, uint64_t *kl,
                           int16_t *lev) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const P o = load_oct(octs, i);
  const Key k = make_key(o.block, o.x, o.y, o.z);
  kh[i] = k.hi;
  kl[i] = k.lo;
  lev[i] = (int16_t)o.level;
}

// computeDepFacesAndEdges: which faces/edges of each octant hang on a
// coarser neighbour
__global__ void k_dep_info(DBlk B, DOcts O, const Oct *__restrict__ octs, int16_t *info) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= O.n) return;
  const P o = load_oct(octs, i);
  if (o.level == 0) {
    info[i] = 0;
    return;
  }
  int face_info = 0, edge_info = 0;
  const int id = child_id(o);
  const P par = parent_of(o);
  for (int k = 0; k < 3; k++) {
    const int fi = cChildFaces[id][k];
    const P nb = face_neighbor(par, fi);
    if (in_bounds(nb)) {
      if (contains_exact(O, nb)) face_info |= 1 << fi;
    } else if (check_adjacent_faces(B, O, fi, nb)) {
      face_info |= 1 << fi;
    }
  }
  for (int k = 0; k < 3; k++) {
    const int ei = cChildEdges[id][k];
    const P nb = edge_neighbor(par, ei);
    const int fx0 = nb.x < 0, fy0 = nb.y < 0, fz0 = nb.z < 0;
    const int fx = fx0 || nb.x >= HMAX, fy = fy0 || nb.y >= HMAX, fz = fz0 || nb.z >= HMAX;
    if ((fx && fy) || (fy && fz) || (fx && fz)) {
      if (check_adjacent_edges(B, O, ei, nb)) edge_info |= 1 << ei;
    } else if (fx || fy || fz) {
      const int fi = fx * (fx0 ? 0 : 1) + fy * (fy0 ? 2 : 3) + fz * (fz0 ? 4 : 5);
      if (check_adjacent_faces(B, O, fi, nb)) edge_info |= 1 << ei;
    } else if (contains_exact(O, nb)) {
      edge_info |= 1 << ei;
    }
  }
  int v = 0;
  for (int k = 0; k < 3; k++) {
    if (face_info & (1 << cChildFaces[id][k])) v |= 1 << k;
    if (edge_info & (1 << cChildEdges[id][k])) v |= 1 << (k + 3);
  }
  info[i] = (int16_t)v;
}

__global__ void k_corner_keys(DBlk B, int n, const Oct *__restrict__ octs, uint64_t *kh,
                              uint64_t *kl, uint32_t *ref) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 8 * (int64_t)n) return;
  const int i = (int)(t >> 3), c = (int)(t & 7);
  P nd = corner_node(load_oct(octs, i), c);
  transform_node(B, nd);
  const Key k = make_key(nd.block, nd.x, nd.y, nd.z);
  kh[t] = k.hi;
  kl[t] = k.lo;
  ref[t] = (uint32_t)t;
}

__global__ void k_gather_u64(int64_t m, const uint32_t *__restrict__ idx,
                             const uint64_t *__restrict__ src, uint64_t *dst) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j < m) dst[j] = src[idx[j]];
}

__global__ void k_run_starts(int64_t m, const uint64_t *__restrict__ kh,
                             const uint64_t *__restrict__ kl, int *flag) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= m) return;
  flag[j] = (j == 0 || kh[j] != kh[j - 1] || kl[j] != kl[j - 1]) ? 1 : 0;
}

// conn in node order; the unique keys and node positions
__global__ void k_nodes(DBlk B, int64_t m, const Oct *__restrict__ octs,
                        const uint32_t *__restrict__ ref, const int *__restrict__ flag,
                        const int *__restrict__ incl, const uint64_t *__restrict__ kh,
                        const uint64_t *__restrict__ kl, int *conn, uint64_t *nkh, uint64_t *nkl,
                        int4 *npos) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= m) return;
  const int id = incl[j] - 1;
  const uint32_t r = ref[j];
  conn[r] = id;
  if (flag[j]) {
    nkh[id] = kh[j];
    nkl[id] = kl[j];
    P nd = corner_node(load_oct(octs, (int)(r >> 3)), (int)(r & 7));
    transform_node(B, nd);
    npos[id] = make_int4(nd.block, nd.x, nd.y, nd.z);
  }
}

// every node a hanging constraint refers to must be a node of the mesh
__global__ void k_check_parents(DBlk B, DNodes N, int n, const Oct *__restrict__ octs,
                                const int16_t *__restrict__ info, int *bad) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || !info[i]) return;
  const P o = load_oct(octs, i);
  int fi, ei;
  decode_info(child_id(o), info[i], fi, ei);
  const P par = parent_of(o);
  for (int e = 0; e < 12; e++) {
    if (!(ei & (1 << e))) continue;
    for (int ii = 0; ii < 2; ii++) {
      P nd = parent_edge_node(par, e, ii);
      transform_node(B, nd);
      if (find_node(N, nd) < 0) atomicOr(bad, 1);
    }
  }
}

// labelDependentNodes: the corner at the middle of each hanging parent edge
// (the edges of hanging faces included)
__global__ void k_label_dep(int n, const Oct *__restrict__ octs, const int16_t *__restrict__ info,
                            const int *__restrict__ conn, int *isdep) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || !info[i]) return;
  const P o = load_oct(octs, i);
  const int id0 = child_id(o);
  int fi, ei;
  decode_info(id0, info[i], fi, ei);
  for (int f = 0; f < 6; f++)
    if (fi & (1 << f))
      for (int k = 0; k < 4; k++) ei |= 1 << cFaceEdges[f][k];
  for (int e = 0; e < 12; e++) {
    if (!(ei & (1 << e))) continue;
    const int id = (e < 4) ? id0 % 2 : (e < 8) ? (id0 % 4) / 2 : id0 / 4;
    const int s = (id == 0) ? 1 : 0;
    int off;
    if (e < 4)
      off = s + 2 * (e % 2) + 4 * (e / 2);
    else if (e < 8)
      off = (e % 2) + 2 * s + 4 * ((e - 4) / 2);
    else
      off = (e % 2) + 2 * ((e - 8) / 2) + 4 * s;
    isdep[conn[8 * (size_t)i + off]] = 1;
  }
}

// constraint sizes: 2 for edge-hanging nodes, then 4 for those still unset
__global__ void k_dep_size(int n, int pass, const Oct *__restrict__ octs,
                           const int16_t *__restrict__ info, const int *__restrict__ conn,
                           const int *__restrict__ isdep, const int *__restrict__ dincl, int *dsz) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || !info[i]) return;
  const P o = load_oct(octs, i);
  int fi, ei;
  decode_info(child_id(o), info[i], fi, ei);
  const int *c = conn + 8 * (size_t)i;
  if (pass == 0) {
    for (int e = 0; e < 12; e++) {
      if (!(ei & (1 << e))) continue;
      for (int k = 0; k < 2; k++) {
        const int nd = c[edge_off(e, k)];
        if (isdep[nd]) dsz[dincl[nd]] = 2;  // dependent index + 1
      }
    }
  } else {
    for (int f = 0; f < 6; f++) {
      if (!(fi & (1 << f))) continue;
      for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++) {
          const int nd = c[face_off(f, a, b)];
          if (isdep[nd]) atomicCAS(&dsz[dincl[nd]], 0, 4);
        }
    }
  }
}

// createDependentConn: each octant posts its rank for the constraints it
// would write (WRITE false), then the highest rank writes them (WRITE true)
template <bool WRITE>
__global__ void k_dep_conn(DBlk B, DNodes N, int n, const Oct *__restrict__ octs,
                           const int16_t *__restrict__ info, const int *__restrict__ conn,
                           const int *__restrict__ isdep, const int *__restrict__ dincl,
                           const int *__restrict__ dptr, int *rank, int *dconn, double *dw) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || !info[i]) return;
  const P o = load_oct(octs, i);
  const int id = child_id(o);
  int fi, ei;
  decode_info(id, info[i], fi, ei);
  const P par = parent_of(o);
  const int bx = id % 2, by = (id % 4) / 2, bz = id / 4;
  const int *c = conn + 8 * (size_t)i;
  const double knots[2] = {-1.0, 1.0};
  for (int e = 0; e < 12; e++) {
    if (!(ei & (1 << e))) continue;
    int en[2] = {-1, -1};
    if (WRITE)
      for (int ii = 0; ii < 2; ii++) {
        P nd = parent_edge_node(par, e, ii);
        transform_node(B, nd);
        en[ii] = find_node(N, nd);
      }
    for (int k = 0; k < 2; k++) {
      const int nd = c[edge_off(e, k)];
      if (!isdep[nd]) continue;
      const int d = dincl[nd] - 1;
      if (dptr[d + 1] - dptr[d] != 2) continue;
      const int r = 64 * i + 2 * e + k;
      if (!WRITE) {
        atomicMax(&rank[d], r);
        continue;
      }
      if (rank[d] != r) continue;
      const int ax = (e < 4) ? bx : (e < 8) ? by : bz;
      const double u = 1.0 * (ax - 1) + 0.5 * (1.0 + knots[k]);
      dconn[dptr[d]] = en[0];
      dconn[dptr[d] + 1] = en[1];
      dw[dptr[d]] = (u - 1.0) * (1.0 / (-1.0 - 1.0));
      dw[dptr[d] + 1] = (u + 1.0) * (1.0 / (1.0 + 1.0));
    }
  }
  for (int f = 0; f < 6; f++) {
    if (!(fi & (1 << f))) continue;
    int fn[4] = {-1, -1, -1, -1};
    if (WRITE)
      for (int jj = 0; jj < 2; jj++)
        for (int ii = 0; ii < 2; ii++) {
          P nd = parent_face_node(par, f, ii, jj);
          transform_node(B, nd);
          fn[ii + 2 * jj] = find_node(N, nd);
        }
    for (int jj = 0; jj < 2; jj++)
      for (int ii = 0; ii < 2; ii++) {
        const int nd = c[face_off(f, ii, jj)];
        if (!isdep[nd]) continue;
        const int d = dincl[nd] - 1;
        if (dptr[d + 1] - dptr[d] != 4) continue;
        const int r = 64 * i + 24 + 4 * f + 2 * jj + ii;
        if (!WRITE) {
          atomicMax(&rank[d], r);
          continue;
        }
        if (rank[d] != r) continue;
        double u = -1.0 + 0.5 * (1.0 + knots[ii]);
        double v = -1.0 + 0.5 * (1.0 + knots[jj]);
        if (f < 2) {
          u += 1.0 * by; v += 1.0 * bz;
        } else if (f < 4) {
          u += 1.0 * bx; v += 1.0 * bz;
        } else {
          u += 1.0 * bx; v += 1.0 * by;
        }
        const double Nu[2] = {(u - 1.0) * (1.0 / (-1.0 - 1.0)), (u + 1.0) * (1.0 / (1.0 + 1.0))};
        const double Nv[2] = {(v - 1.0) * (1.0 / (-1.0 - 1.0)), (v + 1.0) * (1.0 / (1.0 + 1.0))};
        for (int j = 0; j < 4; j++) {
          dconn[dptr[d] + j] = fn[j];
          dw[dptr[d] + j] = Nu[j % 2] * Nv[j / 2];
        }
      }
  }
}

// final numbers: independent nodes in node order, dependents -1, -2, ...
__global__ void k_numbers(int nn, const int *__restrict__ isdep, const int *__restrict__ dincl,
                          int *nums) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nn) return;
  // dincl counts dependents up to j, so j + 1 - dincl counts independents
  nums[j] = isdep[j] ? -dincl[j] : (j - dincl[j]);
}

__global__ void k_renumber(int64_t m, const int *__restrict__ nums, int *v) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j < m) v[j] = nums[v[j]];
}

// node locations: the first octant corner that names a node sets its point
__global__ void k_first_corner(int n, int ndep, const int *__restrict__ conn, int *first) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 8 * (int64_t)n) return;
  const int node = conn[t];
  const int index = node >= 0 ? node + ndep : ndep - 1 - (-node - 1);
  atomicMin(&first[index], (int)t);
}

__device__ __forceinline__ double to_coordinate(int32_t x) {
#pragma clang fp contract(off)
  if (x == 0) return 0.0;
  if (x == HMAX - 1) return 1.0;
  return 1.0 * x / HMAX;
}

__global__ void k_points(int nloc, const Oct *__restrict__ octs, const int *__restrict__ first,
                         const int *__restrict__ block_conn, const double *__restrict__ Xv,
                         double *X) {
#pragma clang fp contract(off)
  const int index = blockIdx.x * blockDim.x + threadIdx.x;
  if (index >= nloc) return;
  const int t = first[index], i = t >> 3, c = t & 7;
  const P o = load_oct(octs, i);
  const double d = to_coordinate(hsize(o.level));
  const double u0 = to_coordinate(o.x), v0 = to_coordinate(o.y), w0 = to_coordinate(o.z);
  const double knots[2] = {-1.0, 1.0};
  const double u = u0 + 0.5 * d * (1.0 + knots[c & 1]);
  const double v = v0 + 0.5 * d * (1.0 + knots[(c >> 1) & 1]);
  const double w = w0 + 0.5 * d * (1.0 + knots[c >> 2]);
  const int *bc = block_conn + 8 * o.block;
  double p[3] = {0.0, 0.0, 0.0};
  for (int k = 0; k < 8; k++) {
    const double N = ((k & 1) ? u : 1.0 - u) * ((k & 2) ? v : 1.0 - v) * ((k & 4) ? w : 1.0 - w);
    for (int a = 0; a < 3; a++) p[a] += N * Xv[3 * bc[k] + a];
  }
  for (int a = 0; a < 3; a++) X[3 * (size_t)index + a] = p[a];
}

// --------------------------------------------------------------------------
// createInterpolation (src/TMROctForest.cpp:6396-6793): one row per fine
// independent node, from the first element corner (in element order) that
// names it; the coarse octant enclosing that corner (findEnclosing,
// :6228-6377), its trilinear weights there, coarse hanging nodes expanded,
// then sorted by column (stable) and merged.
__device__ __forceinline__ int cmp_pos(const Key &a, const Key &b) {
  return key_less(a, b) ? -1 : (key_less(b, a) ? 1 : 0);
}

__device__ int find_enclosing(int size, const Oct *__restrict__ co, const P &node, int info) {
  const int32_t h = hsize(node.level);
  const int ii = info % 2, jj = (info % 4) / 2, kk = info / 4;
  const int32_t xi = node.x + ii * h, yi = node.y + jj * h, zi = node.z + kk * h;
  const Key kn = make_key(node.block, node.x, node.y, node.z);
  auto contains = [&](const Oct &a) {
    const int32_t ha = hsize(a.level);
    return node.block == a.block && node.x >= a.x && node.x < a.x + ha && node.y >= a.y &&
           node.y < a.y + ha && node.z >= a.z && node.z < a.z + ha;
  };
  int low = 0, high = size - 1;
  int mid = low + (high - low) / 2;
  while (mid != low) {
    const Oct a = co[mid];
    if (contains(a)) break;
    const int stat = cmp_pos(make_key(a.block, a.x, a.y, a.z), kn);
    if (stat == 0) break;
    if (stat < 0)
      low = mid + 1;
    else
      high = mid - 1;
    mid = low + (high - low) / 2;
  }
  const Key bound = make_key(node.block, node.x + h, node.y + h, node.z + h);
  while (mid < size) {
    const Oct a = co[mid];
    if (cmp_pos(make_key(a.block, a.x, a.y, a.z), bound) > 0) break;
    if (a.block == node.block) {
      const int32_t hm = hsize(a.level);
      if (a.x <= xi && xi <= a.x + hm && a.y <= yi && yi <= a.y + hm && a.z <= zi &&
          zi <= a.z + hm)
        return mid;
    }
    mid++;
  }
  return -1;
}

struct DInterp {
  int nf;  // fine octants
  const Oct *fo;
  const int *fconn;
  int nc;  // coarse octants
  const Oct *co;
  const int *cconn, *cptr, *cdc;
  const double *cdw;
};

// the first corner naming each fine independent node
__global__ void k_first_visit(int64_t m, const int *__restrict__ conn, int *first) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= m) return;
  const int r = conn[t];
  if (r >= 0) atomicMin(&first[r], (int)t);
}

// row of fine node r: length (WRITE false) or entries at rowp[r]
template <bool WRITE>
__global__ void k_interp_rows(DInterp D, int n, const int *__restrict__ first, int *len,
                              const int *__restrict__ rowp, int *cols, double *w, int *bad) {
#pragma clang fp contract(off)
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const int t0 = first[r], i = t0 >> 3, j = t0 & 7;
  const P node = load_oct(D.fo, i);
  const int tc = find_enclosing(D.nc, D.co, node, j);
  if (tc < 0) {
    atomicOr(bad, 1);
    if (!WRITE) len[r] = 0;
    return;
  }
  const Oct t = D.co[tc];
  const int ijk[3] = {j % 2, (j % 4) / 2, j / 4};
  const int32_t h = hsize(node.level), hc = hsize(t.level);
  const int32_t nx[3] = {node.x, node.y, node.z}, tx[3] = {t.x, t.y, t.z};
  const double knots[2] = {-1.0, 1.0};
  double N[3][2] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
  int st[3], en[3];
  for (int a = 0; a < 3; a++) {
    st[a] = 0;
    en[a] = 2;
    if ((ijk[a] == 0 && tx[a] == nx[a]) || (ijk[a] == 1 && tx[a] == nx[a] + h)) {
      st[a] = 0; en[a] = 1; N[a][0] = 1.0;
    } else if ((ijk[a] == 0 && tx[a] + hc == nx[a]) || (ijk[a] == 1 && tx[a] + hc == nx[a] + h)) {
      st[a] = 1; en[a] = 2; N[a][1] = 1.0;
    } else {
      const double u = -1.0 + 2.0 * (nx[a] + 0.5 * h * (1.0 + knots[ijk[a]]) - tx[a]) / hc;
      N[a][0] = (u - 1.0) * (1.0 / (-1.0 - 1.0));
      N[a][1] = (u + 1.0) * (1.0 / (1.0 + 1.0));
    }
  }
  // at most 8 corners, each expanded into at most 4 parents
  int kc[32];
  double kw[32];
  int nk = 0;
  const int *cc = D.cconn + 8 * (size_t)t.tag;
  for (int z = st[2]; z < en[2]; z++)
    for (int y = st[1]; y < en[1]; y++)
      for (int x = st[0]; x < en[0]; x++) {
        const double wv = N[0][x] * N[1][y] * N[2][z];
        const int cn = cc[x + 2 * y + 4 * z];
        if (cn >= 0) {
          kc[nk] = cn;
          kw[nk++] = wv;
        } else {
          const int d = -cn - 1;
          for (int jp = D.cptr[d]; jp < D.cptr[d + 1] && nk < 32; jp++) {
            kc[nk] = D.cdc[jp];
            kw[nk++] = wv * D.cdw[jp];
          }
        }
      }
  // stable insertion sort by column, then merge equal columns in order
  for (int a = 1; a < nk; a++) {
    const int c = kc[a];
    const double v = kw[a];
    int b = a - 1;
    while (b >= 0 && kc[b] > c) {
      kc[b + 1] = kc[b];
      kw[b + 1] = kw[b];
      b--;
    }
    kc[b + 1] = c;
    kw[b + 1] = v;
  }
  int out = 0;
  for (int a = 0; a < nk; a++) {
    if (out > 0 && kc[out - 1] == kc[a]) {
      kw[out - 1] += kw[a];
    } else {
      kc[out] = kc[a];
      kw[out++] = kw[a];
    }
  }
  if (!WRITE) {
    len[r] = out;
    return;
  }
  const int base = rowp[r];
  for (int a = 0; a < out; a++) {
    cols[base + a] = kc[a];
    w[base + a] = kw[a];
  }
}

// --------------------------------------------------------------------------
// balance (src/TMROctForest.cpp:2525-3089) as the host's level sweep: the
// 0-children of every level, sorted and de-duplicated, each adding the
// 0-children of its parent's face, edge (and corner) neighbours one level
// coarser (balanceOctant, across block faces, edges and corners through the
// block connectivity); then all siblings of every level, sorted by position
// and level, keeping the finest octant per position.
__device__ __forceinline__ P sibling0(const P &o) {
  const int32_t h = hsize(o.level);
  P s = o;
  s.x = (o.x & h) ? o.x - h : o.x;
  s.y = (o.y & h) ? o.y - h : o.y;
  s.z = (o.z & h) ? o.z - h : o.z;
  return s;
}

__device__ __forceinline__ P sibling_of(const P &o, int id) {
  const P s = sibling0(o);
  const int32_t h = hsize(o.level);
  return P{o.block, (id & 1) ? s.x + h : s.x, (id & 2) ? s.y + h : s.y, (id & 4) ? s.z + h : s.z,
           o.level};
}

__device__ __forceinline__ P corner_neighbor(const P &o, int corner) {
  const int32_t h = hsize(o.level);
  P n = o;
  n.x = o.x + (2 * (corner & 1) - 1) * h;
  n.y = o.y + ((corner & 2) - 1) * h;
  n.z = o.z + ((corner & 4) / 2 - 1) * h;
  return n;
}

// emits into out[base + k] (when out is not null) and counts
struct Emit {
  Oct *out;
  int base, k;
  __device__ void operator()(const P &q) {
    if (out) out[base + k] = Oct{q.block, q.x, q.y, q.z, 0, (int16_t)q.level, 0};
    k++;
  }
};

__device__ void add_face_neighbors(const DBlk &B, int face_index, const P &p, Emit &em) {
  const int block = p.block;
  const int face = B.block_face_conn[6 * block + face_index];
  const int face_id = B.block_face_ids[6 * block + face_index];
  const int32_t h = hsize(p.level);
  int32_t u, v;
  if (face_index < 2)
    get_face_coords(face_id, HMAX - 2 * h, p.y, p.z, u, v);
  else if (face_index < 4)
    get_face_coords(face_id, HMAX - 2 * h, p.x, p.z, u, v);
  else
    get_face_coords(face_id, HMAX - 2 * h, p.x, p.y, u, v);
  for (int ip = B.face_block_ptr[face]; ip < B.face_block_ptr[face + 1]; ip++) {
    const int adj = B.face_block_conn[ip] / 6;
    if (adj == block) continue;
    const int ai = B.face_block_conn[ip] % 6;
    const int fid = B.block_face_ids[6 * adj + ai];
    P n{adj, 0, 0, 0, p.level};
    if (ai < 2) {
      n.x = (HMAX - 2 * h) * (ai % 2);
      set_face_coords(fid, HMAX - 2 * h, u, v, n.y, n.z);
    } else if (ai < 4) {
      n.y = (HMAX - 2 * h) * (ai % 2);
      set_face_coords(fid, HMAX - 2 * h, u, v, n.x, n.z);
    } else {
      n.z = (HMAX - 2 * h) * (ai % 2);
      set_face_coords(fid, HMAX - 2 * h, u, v, n.x, n.y);
    }
    em(n);
  }
}

__device__ void add_edge_neighbors(const DBlk &B, int edge_index, const P &p, Emit &em) {
  const int block = p.block;
  const int edge = B.block_edge_conn[12 * block + edge_index];
  const int32_t h = hsize(p.level);
  const int32_t ucoord = (edge_index < 4) ? p.x : (edge_index < 8) ? p.y : p.z;
  const int n1 = B.block_conn[8 * block + cEdgeNodes[edge_index][0]];
  const int n2 = B.block_conn[8 * block + cEdgeNodes[edge_index][1]];
  for (int ip = B.edge_block_ptr[edge]; ip < B.edge_block_ptr[edge + 1]; ip++) {
    const int adj = B.edge_block_conn[ip] / 12;
    if (adj == block) continue;
    const int ai = B.edge_block_conn[ip] % 12;
    const int m1 = B.block_conn[8 * adj + cEdgeNodes[ai][0]];
    const int m2 = B.block_conn[8 * adj + cEdgeNodes[ai][1]];
    const int32_t u = (n1 == m2 && n2 == m1) ? HMAX - 2 * h - ucoord : ucoord;
    P n{adj, 0, 0, 0, p.level};
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
    em(n);
  }
}

__device__ void add_corner_neighbors(const DBlk &B, int corner, const P &p, Emit &em) {
  const int block = p.block;
  const int node = B.block_conn[8 * block + corner];
  const int32_t h = hsize(p.level);
  for (int ip = B.node_block_ptr[node]; ip < B.node_block_ptr[node + 1]; ip++) {
    const int adj = B.node_block_conn[ip] / 8;
    if (adj == block) continue;
    const int ai = B.node_block_conn[ip] % 8;
    em(P{adj, (HMAX - 2 * h) * (ai % 2), (HMAX - 2 * h) * ((ai % 4) / 2),
         (HMAX - 2 * h) * (ai / 4), p.level});
  }
}

__device__ void balance_octant(const DBlk &B, const P &oct, int balance_corner, Emit &em) {
  if (oct.level <= 1) return;
  const P p = parent_of(oct);
  for (int f = 0; f < 6; f++) {
    const P q = sibling0(face_neighbor(p, f));
    if (in_bounds(q))
      em(q);
    else
      add_face_neighbors(B, f, q, em);
  }
  for (int e = 0; e < 12; e++) {
    const P q = sibling0(edge_neighbor(p, e));
    if (in_bounds(q)) {
      em(q);
      continue;
    }
    const int ex = (q.x < 0 || q.x >= HMAX), ey = (q.y < 0 || q.y >= HMAX),
              ez = (q.z < 0 || q.z >= HMAX);
    if ((ex && ey) || (ex && ez) || (ey && ez)) {
      add_edge_neighbors(B, e, q, em);
    } else {
      const int face = ex ? (q.x < 0 ? 0 : 1) : ey ? (q.y < 0 ? 2 : 3) : (q.z < 0 ? 4 : 5);
      add_face_neighbors(B, face, q, em);
    }
  }
  if (!balance_corner) return;
  for (int c = 0; c < 8; c++) {
    const P q = sibling0(corner_neighbor(p, c));
    if (in_bounds(q)) {
      em(q);
      continue;
    }
    const int ex = (q.x < 0 || q.x >= HMAX), ey = (q.y < 0 || q.y >= HMAX),
              ez = (q.z < 0 || q.z >= HMAX);
    if (ex && ey && ez) {
      add_corner_neighbors(B, c, q, em);
    } else if ((ex && ey) || (ex && ez) || (ey && ez)) {
      int edge;
      if (ey && ez)
        edge = (q.y < 0 ? 0 : 1) + (q.z < 0 ? 0 : 2);
      else if (ex && ez)
        edge = (q.x < 0 ? 4 : 5) + (q.z < 0 ? 0 : 2);
      else
        edge = (q.x < 0 ? 8 : 9) + (q.y < 0 ? 0 : 2);
      add_edge_neighbors(B, edge, q, em);
    } else {
      const int face = ex ? (q.x < 0 ? 0 : 1) : ey ? (q.y < 0 ? 2 : 3) : (q.z < 0 ? 4 : 5);
      add_face_neighbors(B, face, q, em);
    }
  }
}

// the 0-child of every octant (tag and info cleared)
__global__ void k_sibling0(int n, const Oct *__restrict__ in, Oct *out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const P s = sibling0(load_oct(in, i));
  out[i] = Oct{s.block, s.x, s.y, s.z, 0, (int16_t)s.level, 0};
}

// balanceOctant's additions: counts (out null) or the octants at offs[i]
__global__ void k_balance_emit(DBlk B, int n, const Oct *__restrict__ in, int corner,
                               const int *__restrict__ offs, int *cnt, Oct *out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  Emit em{out, out ? offs[i] : 0, 0};
  balance_octant(B, load_oct(in, i), corner, em);
  if (!out) cnt[i] = em.k;
}

// the 8 siblings of each octant (of one level > 0)
__global__ void k_siblings(int n, const Oct *__restrict__ in, Oct *out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const P o = load_oct(in, i);
  for (int j = 0; j < 8; j++) {
    const P s = sibling_of(o, j);
    out[8 * (size_t)i + j] = Oct{s.block, s.x, s.y, s.z, 0, (int16_t)s.level, 0};
  }
}

__global__ void k_sort_keys(int n, const Oct *__restrict__ a, uint64_t *kh, uint64_t *kl,
                            uint64_t *klev, uint32_t *idx) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const Oct o = a[i];
  const Key k = make_key(o.block, o.x, o.y, o.z);
  kh[i] = k.hi;
  kl[i] = k.lo;
  if (klev) klev[i] = (uint64_t)o.level;
  idx[i] = (uint32_t)i;
}

__global__ void k_gather_oct(int n, const uint32_t *__restrict__ idx, const Oct *__restrict__ a,
                             Oct *b) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = a[idx[i]];
}

// keep[i]: first (last) entry of each run of equal keys
__global__ void k_run_flags(int n, const uint64_t *__restrict__ kh, const uint64_t *__restrict__ kl,
                            int last, int *keep) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int j = last ? i + 1 : i - 1;
  const bool edge = last ? (j >= n) : (j < 0);
  keep[i] = (edge || kh[j] != kh[i] || kl[j] != kl[i]) ? 1 : 0;
}

__global__ void k_compact_oct(int n, const int *__restrict__ keep, const int *__restrict__ pos,
                              const Oct *__restrict__ a, Oct *b, int set_tag) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || !keep[i]) return;
  const int j = pos[i] - 1;
  Oct o = a[i];
  if (set_tag) {
    o.tag = j;
    o.info = 0;
  }
  b[j] = o;
}

__global__ void k_excl(int n, const int *__restrict__ c, const int *__restrict__ ci, int *co) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) co[i] = ci[i] - c[i];
}

__global__ void k_key_words(int n, const Oct *__restrict__ a, uint64_t *kh, uint64_t *kl) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const Key k = make_key(a[i].block, a[i].x, a[i].y, a[i].z);
  kh[i] = k.hi;
  kl[i] = k.lo;
}

// --------------------------------------------------------------------------
// device buffers released on every exit
struct Pool {
  std::vector<void *> p;
  ~Pool() {
    for (void *q : p) (void)hipFree(q);
  }
  template <class T>
  T *get(size_t n) {
    void *q = nullptr;
    if (hipMalloc(&q, (n ? n : 1) * sizeof(T)) != hipSuccess) return nullptr;
    p.push_back(q);
    return (T *)q;
  }
  template <class T>
  T *up(const std::vector<T> &v) {
    T *q = get<T>(v.size());
    if (q && !v.empty() && hipMemcpy(q, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice) != hipSuccess)
      return nullptr;
    return q;
  }
};

inline unsigned blocks(int64_t n, int b = 256) { return (unsigned)((n + b - 1) / b); }

}  // namespace

int Forest::create_nodes_device(int device, double *ms) {
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
    u128 vol = 0;
    for (const Oct &o : octs_) vol += (u128)1 << (3 * (MAX_LEVEL - o.level));
    if (vol != (u128)bd_.num_blocks << (3 * MAX_LEVEL)) {
      last_error = "create_nodes: the octants do not tile the blocks; call balance() first";
      return -1;
    }
  }
  if ((int64_t)octs_.size() * 64 >= (int64_t)1 << 31) {
    last_error = "create_nodes_device: more than 2^25 octants";
    return -1;
  }
  // restores the caller's device after the buffers below are freed
  struct DeviceGuard {
    int prev = 0;
    DeviceGuard() { (void)hipGetDevice(&prev); }
    ~DeviceGuard() { (void)hipSetDevice(prev); }
  } guard;
  if (hipSetDevice(device) != hipSuccess) {
    last_error = "create_nodes_device: cannot select device " + std::to_string(device);
    return -1;
  }
  auto t0 = std::chrono::steady_clock::now();
  {
    Pool pool;
    const int n = (int)octs_.size();
    const int64_t m = 8 * (int64_t)n;
    const BlockConn &bc = bd_;
    auto fail = [&](const char *what) {
      last_error = std::string("create_nodes_device: ") + what;
      return -1;
    };
    DBlk B;
    B.block_conn = pool.up(bc.block_conn);
    B.block_edge_conn = pool.up(bc.block_edge_conn);
    B.block_face_conn = pool.up(bc.block_face_conn);
    B.block_face_ids = pool.up(bc.block_face_ids);
    B.node_block_ptr = pool.up(bc.node_block_ptr);
    B.node_block_conn = pool.up(bc.node_block_conn);
    B.edge_block_ptr = pool.up(bc.edge_block_ptr);
    B.edge_block_conn = pool.up(bc.edge_block_conn);
    B.face_block_ptr = pool.up(bc.face_block_ptr);
    B.face_block_conn = pool.up(bc.face_block_conn);
    B.face_owner = pool.up(bc.face_block_owners);
    B.edge_owner = pool.up(bc.edge_block_owners);
    B.node_owner = pool.up(bc.node_block_owners);
    Oct *d_octs = pool.up(octs_);
    uint64_t *okh = pool.get<uint64_t>(n), *okl = pool.get<uint64_t>(n);
    int16_t *olev = pool.get<int16_t>(n), *info = pool.get<int16_t>(n);
    uint64_t *kh = pool.get<uint64_t>(m), *kl = pool.get<uint64_t>(m);
    uint64_t *kh2 = pool.get<uint64_t>(m), *kl2 = pool.get<uint64_t>(m);
    uint32_t *ref = pool.get<uint32_t>(m), *ref2 = pool.get<uint32_t>(m);
    int *flag = pool.get<int>(m), *incl = pool.get<int>(m), *conn = pool.get<int>(m);
    int *bad = pool.get<int>(1);
    if (!B.block_conn || !B.block_edge_conn || !B.block_face_conn || !B.block_face_ids ||
        !B.node_block_ptr || !B.node_block_conn || !B.edge_block_ptr || !B.edge_block_conn ||
        !B.face_block_ptr || !B.face_block_conn || !B.face_owner || !B.edge_owner ||
        !B.node_owner || !d_octs || !okh || !okl || !olev || !info || !kh || !kl || !kh2 ||
        !kl2 || !ref || !ref2 || !flag || !incl || !conn || !bad)
      return fail("device allocation failed");
    // scratch of the radix sorts and scans (the largest request)
    size_t tb = 0, tq = 0;
    (void)hipcub::DeviceRadixSort::SortPairs(nullptr, tq, kl, kl2, ref, ref2, (int)m);
    tb = tq;
    (void)hipcub::DeviceScan::InclusiveSum(nullptr, tq, flag, incl, (int)m);
    tb = std::max(tb, tq);
    void *tmp = pool.get<unsigned char>(tb);
    if (!tmp) return fail("device allocation failed");

    hipLaunchKernelGGL(k_oct_keys, dim3(blocks(n)), dim3(256), 0, 0, n, d_octs, okh, okl, olev);
    DOcts O{n, okh, okl, olev};
    hipLaunchKernelGGL(k_dep_info, dim3(blocks(n)), dim3(256), 0, 0, B, O, d_octs, info);
    hipLaunchKernelGGL(k_corner_keys, dim3(blocks(m)), dim3(256), 0, 0, B, n, d_octs, kh, kl, ref);
    // two stable radix passes give the (high, low) order: by the low word
    // (permutation ref2), then by the high words taken in that order (kl is
    // reused to hold them); the final permutation lands in ref, the sorted
    // high words in kh2
    tq = tb;
    if (hipcub::DeviceRadixSort::SortPairs(tmp, tq, kl, kl2, ref, ref2, (int)m) != hipSuccess)
      return fail("radix sort failed");
    hipLaunchKernelGGL(k_gather_u64, dim3(blocks(m)), dim3(256), 0, 0, m, ref2, kh, kl);
    tq = tb;
    if (hipcub::DeviceRadixSort::SortPairs(tmp, tq, kl, kh2, ref2, ref, (int)m) != hipSuccess)
      return fail("radix sort failed");
    // the corner keys again (kh, kl; ref2 becomes the identity, unused), and
    // the low words in sorted order into kl2
    hipLaunchKernelGGL(k_corner_keys, dim3(blocks(m)), dim3(256), 0, 0, B, n, d_octs, kh, kl, ref2);
    hipLaunchKernelGGL(k_gather_u64, dim3(blocks(m)), dim3(256), 0, 0, m, ref, kl, kl2);
    hipLaunchKernelGGL(k_run_starts, dim3(blocks(m)), dim3(256), 0, 0, m, kh2, kl2, flag);
    tq = tb;
    if (hipcub::DeviceScan::InclusiveSum(tmp, tq, flag, incl, (int)m) != hipSuccess)
      return fail("scan failed");
    int nn = 0;
    if (hipMemcpy(&nn, incl + (m - 1), sizeof(int), hipMemcpyDeviceToHost) != hipSuccess)
      return fail("copy failed");
    uint64_t *nkh = pool.get<uint64_t>(nn), *nkl = pool.get<uint64_t>(nn);
    int4 *npos = pool.get<int4>(nn);
    int *isdep = pool.get<int>(nn), *dincl = pool.get<int>(nn), *nums = pool.get<int>(nn);
    if (!nkh || !nkl || !npos || !isdep || !dincl || !nums)
      return fail("device allocation failed");
    hipLaunchKernelGGL(k_nodes, dim3(blocks(m)), dim3(256), 0, 0, B, m, d_octs, ref, flag, incl,
                       kh2, kl2, conn, nkh, nkl, npos);
    DNodes N{nn, nkh, nkl};
    (void)hipMemset(bad, 0, sizeof(int));
    hipLaunchKernelGGL(k_check_parents, dim3(blocks(n)), dim3(256), 0, 0, B, N, n, d_octs, info,
                       bad);
    int hbad = 0;
    if (hipMemcpy(&hbad, bad, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess)
      return fail("copy failed");
    if (hbad) {
      last_error = "create_nodes: forest is not 2:1 balanced (missing parent node)";
      free_mesh();
      return -1;
    }
    (void)hipMemset(isdep, 0, nn * sizeof(int));
    hipLaunchKernelGGL(k_label_dep, dim3(blocks(n)), dim3(256), 0, 0, n, d_octs, info, conn, isdep);
    size_t tq2 = 0;
    (void)hipcub::DeviceScan::InclusiveSum(nullptr, tq2, isdep, dincl, nn);
    if (tq2 > tb) return fail("scan scratch too small");
    tq2 = tb;
    if (hipcub::DeviceScan::InclusiveSum(tmp, tq2, isdep, dincl, nn) != hipSuccess)
      return fail("scan failed");
    int ndep = 0;
    if (hipMemcpy(&ndep, dincl + (nn - 1), sizeof(int), hipMemcpyDeviceToHost) != hipSuccess)
      return fail("copy failed");
    int *dsz = pool.get<int>(ndep + 1), *dptr = pool.get<int>(ndep + 1);
    int *rank = pool.get<int>(ndep);
    if (!dsz || !dptr || !rank) return fail("device allocation failed");
    std::vector<int> hdptr(ndep + 1, 0);
    int ntot = 0;
    if (ndep > 0) {
      (void)hipMemset(dsz, 0, (ndep + 1) * sizeof(int));
      hipLaunchKernelGGL(k_dep_size, dim3(blocks(n)), dim3(256), 0, 0, n, 0, d_octs, info, conn,
                         isdep, dincl, dsz);
      hipLaunchKernelGGL(k_dep_size, dim3(blocks(n)), dim3(256), 0, 0, n, 1, d_octs, info, conn,
                         isdep, dincl, dsz);
      tq2 = 0;
      (void)hipcub::DeviceScan::InclusiveSum(nullptr, tq2, dsz, dptr, ndep + 1);
      if (tq2 > tb) return fail("scan scratch too small");
      tq2 = tb;
      if (hipcub::DeviceScan::InclusiveSum(tmp, tq2, dsz, dptr, ndep + 1) != hipSuccess)
        return fail("scan failed");
      if (hipMemcpy(hdptr.data(), dptr, (ndep + 1) * sizeof(int), hipMemcpyDeviceToHost) != hipSuccess)
        return fail("copy failed");
      ntot = hdptr[ndep];
    }
    int *dconn = pool.get<int>(ntot);
    double *dw = pool.get<double>(ntot);
    if (!dconn || !dw) return fail("device allocation failed");
    if (ndep > 0) {
      (void)hipMemset(rank, 0xff, ndep * sizeof(int));  // -1
      hipLaunchKernelGGL((k_dep_conn<false>), dim3(blocks(n)), dim3(256), 0, 0, B, N, n, d_octs,
                         info, conn, isdep, dincl, dptr, rank, dconn, dw);
      hipLaunchKernelGGL((k_dep_conn<true>), dim3(blocks(n)), dim3(256), 0, 0, B, N, n, d_octs,
                         info, conn, isdep, dincl, dptr, rank, dconn, dw);
    }
    hipLaunchKernelGGL(k_numbers, dim3(blocks(nn)), dim3(256), 0, 0, nn, isdep, dincl, nums);
    hipLaunchKernelGGL(k_renumber, dim3(blocks(m)), dim3(256), 0, 0, m, nums, conn);
    if (ntot > 0)
      hipLaunchKernelGGL(k_renumber, dim3(blocks(ntot)), dim3(256), 0, 0, (int64_t)ntot, nums,
                         dconn);
    // node locations
    std::vector<double> X(3 * (size_t)nn, 0.0);
    if (!Xv_.empty()) {
      int *first = pool.get<int>(nn);
      double *dX = pool.get<double>(3 * (size_t)nn);
      double *dXv = pool.up(Xv_);
      if (!first || !dX || !dXv) return fail("device allocation failed");
      (void)hipMemset(first, 0x7f, nn * sizeof(int));
      hipLaunchKernelGGL(k_first_corner, dim3(blocks(m)), dim3(256), 0, 0, n, ndep, conn, first);
      hipLaunchKernelGGL(k_points, dim3(blocks(nn)), dim3(256), 0, 0, nn, d_octs, first,
                         B.block_conn, dXv, dX);
      if (hipMemcpy(X.data(), dX, X.size() * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
        return fail("copy failed");
    }
    if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess)
      return fail("kernel launch failed");
    // results back into the forest
    std::vector<int16_t> hinfo(n);
    std::vector<uint64_t> hkh(nn), hkl(nn);
    std::vector<int4> hpos(nn);
    conn_.resize(m);
    dep_conn_.resize(ntot);
    dep_w_.resize(ntot);
    if (hipMemcpy(hinfo.data(), info, n * sizeof(int16_t), hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(conn_.data(), conn, m * sizeof(int), hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(hkh.data(), nkh, nn * sizeof(uint64_t), hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(hkl.data(), nkl, nn * sizeof(uint64_t), hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(hpos.data(), npos, nn * sizeof(int4), hipMemcpyDeviceToHost) != hipSuccess ||
        (ntot > 0 &&
         (hipMemcpy(dep_conn_.data(), dconn, ntot * sizeof(int), hipMemcpyDeviceToHost) != hipSuccess ||
          hipMemcpy(dep_w_.data(), dw, ntot * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess))) {
      free_mesh();
      return fail("copy failed");
    }
    for (int i = 0; i < n; i++) octs_[i].info = hinfo[i];
    nkey_.resize(nn);
    nodes_.resize(nn);
    for (int j = 0; j < nn; j++) {
      nkey_[j] = ((u128)hkh[j] << 64) | (u128)hkl[j];
      nodes_[j] = Oct{hpos[j].x, hpos[j].y, hpos[j].z, hpos[j].w, 0, 1, 0};
    }
    dep_ptr_ = hdptr;
    num_dep_ = ndep;
    num_owned_ = nn - ndep;
    node_numbers_.resize(nn);
    for (int j = 0; j < nn; j++) node_numbers_[j] = j - ndep;
    X_.swap(X);
    nodes_ready_ = true;
  }
  if (ms) *ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return 0;
}

namespace {

// device scratch for sorting and scanning, released with the pool
struct Dev {
  Pool &pool;
  std::string err;
  // out = inclusive prefix sums of in[0..n)
  bool scan(const int *in, int *out, int n) {
    if (n <= 0) return true;
    size_t tb = 0;
    (void)hipcub::DeviceScan::InclusiveSum(nullptr, tb, in, out, n);
    void *tmp = pool.get<unsigned char>(tb);
    if (!tmp || hipcub::DeviceScan::InclusiveSum(tmp, tb, in, out, n) != hipSuccess) {
      err = "scan failed";
      return false;
    }
    return true;
  }
  int last(const int *a, int n) {
    int v = 0;
    if (n > 0) (void)hipMemcpy(&v, a + (n - 1), sizeof(int), hipMemcpyDeviceToHost);
    return v;
  }
  // b = a sorted by position key (by level first, when by_level: stable
  // passes level, low word, high word); kh/kl receive the sorted key words
  bool sort(int n, const Oct *a, Oct *b, bool by_level, uint64_t **khs, uint64_t **kls) {
    uint64_t *kh = pool.get<uint64_t>(n), *kl = pool.get<uint64_t>(n);
    uint64_t *klev = by_level ? pool.get<uint64_t>(n) : nullptr;
    uint64_t *g = pool.get<uint64_t>(n), *ks = pool.get<uint64_t>(n);
    uint32_t *p1 = pool.get<uint32_t>(n), *p2 = pool.get<uint32_t>(n);
    if (!kh || !kl || (by_level && !klev) || !g || !ks || !p1 || !p2) {
      err = "device allocation failed";
      return false;
    }
    size_t tb = 0;
    (void)hipcub::DeviceRadixSort::SortPairs(nullptr, tb, g, ks, p1, p2, n);
    void *tmp = pool.get<unsigned char>(tb);
    if (!tmp) {
      err = "device allocation failed";
      return false;
    }
    hipLaunchKernelGGL(k_sort_keys, dim3(blocks(n)), dim3(256), 0, 0, n, a, kh, kl, klev, p1);
    auto pass = [&](const uint64_t *key, int end_bit) {
      hipLaunchKernelGGL(k_gather_u64, dim3(blocks(n)), dim3(256), 0, 0, (int64_t)n, p1, key, g);
      size_t t = tb;
      if (hipcub::DeviceRadixSort::SortPairs(tmp, t, g, ks, p1, p2, n, 0, end_bit) != hipSuccess)
        return false;
      std::swap(p1, p2);
      return true;
    };
    if ((by_level && !pass(klev, 8)) || !pass(kl, 64) || !pass(kh, 64)) {
      err = "radix sort failed";
      return false;
    }
    hipLaunchKernelGGL(k_gather_oct, dim3(blocks(n)), dim3(256), 0, 0, n, p1, a, b);
    hipLaunchKernelGGL(k_gather_u64, dim3(blocks(n)), dim3(256), 0, 0, (int64_t)n, p1, kl, g);
    *khs = ks;  // the last pass's sorted high words
    *kls = g;
    return true;
  }
};

}  // namespace

int Forest::balance_device(int balance_corner, int device, double *ms) {
  if (octs_.empty()) {
    last_error = "balance: no octants have been created";
    return -1;
  }
  free_mesh();
  struct DeviceGuard {
    int prev = 0;
    DeviceGuard() { (void)hipGetDevice(&prev); }
    ~DeviceGuard() { (void)hipSetDevice(prev); }
  } guard;
  if (hipSetDevice(device) != hipSuccess) {
    last_error = "balance_device: cannot select device " + std::to_string(device);
    return -1;
  }
  auto t0 = std::chrono::steady_clock::now();
  Pool pool;
  Dev dv{pool, ""};
  auto fail = [&](const std::string &what) {
    last_error = "balance_device: " + what;
    return -1;
  };
  DBlk B;
  B.block_conn = pool.up(bd_.block_conn);
  B.block_edge_conn = pool.up(bd_.block_edge_conn);
  B.block_face_conn = pool.up(bd_.block_face_conn);
  B.block_face_ids = pool.up(bd_.block_face_ids);
  B.node_block_ptr = pool.up(bd_.node_block_ptr);
  B.node_block_conn = pool.up(bd_.node_block_conn);
  B.edge_block_ptr = pool.up(bd_.edge_block_ptr);
  B.edge_block_conn = pool.up(bd_.edge_block_conn);
  B.face_block_ptr = pool.up(bd_.face_block_ptr);
  B.face_block_conn = pool.up(bd_.face_block_conn);
  B.face_owner = pool.up(bd_.face_block_owners);
  B.edge_owner = pool.up(bd_.edge_block_owners);
  B.node_owner = pool.up(bd_.node_block_owners);
  const int n0 = (int)octs_.size();
  Oct *d_in = pool.up(octs_);
  Oct *s0 = pool.get<Oct>(n0), *s0l = pool.get<Oct>(n0);
  if (!B.block_conn || !B.block_edge_conn || !B.block_face_conn || !B.block_face_ids ||
      !B.node_block_ptr || !B.node_block_conn || !B.edge_block_ptr || !B.edge_block_conn ||
      !B.face_block_ptr || !B.face_block_conn || !B.face_owner || !B.edge_owner ||
      !B.node_owner || !d_in || !s0 || !s0l)
    return fail("device allocation failed");
  // the 0-children, grouped by level (a stable pass on the level alone)
  hipLaunchKernelGGL(k_sibling0, dim3(blocks(n0)), dim3(256), 0, 0, n0, d_in, s0);
  std::vector<int> cnt(MAX_LEVEL + 1, 0), off(MAX_LEVEL + 2, 0);
  for (const Oct &o : octs_) cnt[o.level]++;
  for (int L = 0; L <= MAX_LEVEL; L++) off[L + 1] = off[L] + cnt[L];
  {
    uint64_t *klev = pool.get<uint64_t>(n0), *ks = pool.get<uint64_t>(n0);
    uint64_t *kh = pool.get<uint64_t>(n0), *kl = pool.get<uint64_t>(n0);
    uint32_t *p1 = pool.get<uint32_t>(n0), *p2 = pool.get<uint32_t>(n0);
    if (!klev || !ks || !kh || !kl || !p1 || !p2) return fail("device allocation failed");
    size_t tb = 0;
    (void)hipcub::DeviceRadixSort::SortPairs(nullptr, tb, klev, ks, p1, p2, n0, 0, 8);
    void *tmp = pool.get<unsigned char>(tb);
    if (!tmp) return fail("device allocation failed");
    hipLaunchKernelGGL(k_sort_keys, dim3(blocks(n0)), dim3(256), 0, 0, n0, s0, kh, kl, klev, p1);
    if (hipcub::DeviceRadixSort::SortPairs(tmp, tb, klev, ks, p1, p2, n0, 0, 8) != hipSuccess)
      return fail("radix sort failed");
    hipLaunchKernelGGL(k_gather_oct, dim3(blocks(n0)), dim3(256), 0, 0, n0, p2, s0, s0l);
  }
  std::vector<Oct *> uq(MAX_LEVEL + 1, nullptr);
  std::vector<int> nu(MAX_LEVEL + 1, 0);
  Oct *gen = nullptr;
  int ngen = 0;
  for (int L = MAX_LEVEL; L >= 0; L--) {
    const int ncur = cnt[L] + ngen;
    if (ncur == 0) continue;
    Oct *cur = pool.get<Oct>(ncur), *srt = pool.get<Oct>(ncur);
    int *keep = pool.get<int>(ncur), *pos = pool.get<int>(ncur);
    if (!cur || !srt || !keep || !pos) return fail("device allocation failed");
    if (cnt[L])
      (void)hipMemcpy(cur, s0l + off[L], cnt[L] * sizeof(Oct), hipMemcpyDeviceToDevice);
    if (ngen) (void)hipMemcpy(cur + cnt[L], gen, ngen * sizeof(Oct), hipMemcpyDeviceToDevice);
    uint64_t *khs = nullptr, *kls = nullptr;
    if (!dv.sort(ncur, cur, srt, false, &khs, &kls)) return fail(dv.err);
    hipLaunchKernelGGL(k_run_flags, dim3(blocks(ncur)), dim3(256), 0, 0, ncur, khs, kls, 0, keep);
    if (!dv.scan(keep, pos, ncur)) return fail(dv.err);
    nu[L] = dv.last(pos, ncur);
    uq[L] = pool.get<Oct>(nu[L]);
    if (!uq[L]) return fail("device allocation failed");
    hipLaunchKernelGGL(k_compact_oct, dim3(blocks(ncur)), dim3(256), 0, 0, ncur, keep, pos, srt,
                       uq[L], 0);
    ngen = 0;
    gen = nullptr;
    if (L > 1) {
      int *c = pool.get<int>(nu[L]), *ci = pool.get<int>(nu[L]), *co = pool.get<int>(nu[L]);
      if (!c || !ci || !co) return fail("device allocation failed");
      hipLaunchKernelGGL(k_balance_emit, dim3(blocks(nu[L])), dim3(256), 0, 0, B, nu[L], uq[L],
                         balance_corner, nullptr, c, nullptr);
      if (!dv.scan(c, ci, nu[L])) return fail(dv.err);
      ngen = dv.last(ci, nu[L]);
      // exclusive offsets: inclusive sums minus the own count
      hipLaunchKernelGGL(k_excl, dim3(blocks(nu[L])), dim3(256), 0, 0, nu[L], c, ci, co);
      gen = pool.get<Oct>(ngen);
      if (!gen) return fail("device allocation failed");
      if (ngen)
        hipLaunchKernelGGL(k_balance_emit, dim3(blocks(nu[L])), dim3(256), 0, 0, B, nu[L], uq[L],
                           balance_corner, co, nullptr, gen);
    }
  }
  // every sibling, then the finest octant per position
  int nall = 0;
  for (int L = 0; L <= MAX_LEVEL; L++) nall += (L > 0 ? 8 : 1) * nu[L];
  Oct *all = pool.get<Oct>(nall), *srt = pool.get<Oct>(nall);
  int *keep = pool.get<int>(nall), *pos = pool.get<int>(nall);
  if (!all || !srt || !keep || !pos) return fail("device allocation failed");
  {
    size_t at = 0;
    for (int L = 0; L <= MAX_LEVEL; L++) {
      if (!nu[L]) continue;
      if (L == 0) {
        (void)hipMemcpy(all + at, uq[L], nu[L] * sizeof(Oct), hipMemcpyDeviceToDevice);
        at += nu[L];
      } else {
        hipLaunchKernelGGL(k_siblings, dim3(blocks(nu[L])), dim3(256), 0, 0, nu[L], uq[L],
                           all + at);
        at += 8 * (size_t)nu[L];
      }
    }
  }
  uint64_t *khs = nullptr, *kls = nullptr;
  if (!dv.sort(nall, all, srt, true, &khs, &kls)) return fail(dv.err);
  hipLaunchKernelGGL(k_run_flags, dim3(blocks(nall)), dim3(256), 0, 0, nall, khs, kls, 1, keep);
  if (!dv.scan(keep, pos, nall)) return fail(dv.err);
  const int nr = dv.last(pos, nall);
  Oct *res = pool.get<Oct>(nr);
  uint64_t *rh = pool.get<uint64_t>(nr), *rl = pool.get<uint64_t>(nr);
  if (!res || !rh || !rl) return fail("device allocation failed");
  hipLaunchKernelGGL(k_compact_oct, dim3(blocks(nall)), dim3(256), 0, 0, nall, keep, pos, srt, res,
                     1);
  hipLaunchKernelGGL(k_key_words, dim3(blocks(nr)), dim3(256), 0, 0, nr, res, rh, rl);
  if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess)
    return fail("kernel launch failed");
  std::vector<Oct> out(nr);
  std::vector<uint64_t> hh(nr), hl(nr);
  if (hipMemcpy(out.data(), res, nr * sizeof(Oct), hipMemcpyDeviceToHost) != hipSuccess ||
      hipMemcpy(hh.data(), rh, nr * sizeof(uint64_t), hipMemcpyDeviceToHost) != hipSuccess ||
      hipMemcpy(hl.data(), rl, nr * sizeof(uint64_t), hipMemcpyDeviceToHost) != hipSuccess)
    return fail("copy failed");
  octs_.swap(out);
  okey_.resize(nr);
  for (int i = 0; i < nr; i++) okey_[i] = ((u128)hh[i] << 64) | (u128)hl[i];
  if (ms) *ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return 0;
}

int Forest::create_interpolation_device(Forest *coarse, InterpCSR &P, int device, double *ms) {
  if (coarse == this) {
    last_error = "create_interpolation: cannot interpolate between the same forest";
    return -1;
  }
  if (coarse->interp_ != interp_) {
    last_error = "create_interpolation: interpolation types differ";
    return -1;
  }
  if (create_nodes_device(device) != 0) return -1;
  if (coarse->create_nodes_device(device) != 0) {
    last_error = "create_interpolation: coarse forest: " + coarse->last_error;
    return -1;
  }
  struct DeviceGuard {
    int prev = 0;
    DeviceGuard() { (void)hipGetDevice(&prev); }
    ~DeviceGuard() { (void)hipSetDevice(prev); }
  } guard;
  if (hipSetDevice(device) != hipSuccess) {
    last_error = "create_interpolation_device: cannot select device " + std::to_string(device);
    return -1;
  }
  auto t0 = std::chrono::steady_clock::now();
  auto fail = [&](const char *what) {
    last_error = std::string("create_interpolation_device: ") + what;
    return -1;
  };
  Pool pool;
  const int n = num_owned_;
  const int64_t m = 8 * (int64_t)octs_.size();
  DInterp D;
  D.nf = (int)octs_.size();
  D.fo = pool.up(octs_);
  D.fconn = pool.up(conn_);
  D.nc = (int)coarse->octs_.size();
  D.co = pool.up(coarse->octs_);
  D.cconn = pool.up(coarse->conn_);
  D.cptr = pool.up(coarse->dep_ptr_.empty() ? std::vector<int>{0} : coarse->dep_ptr_);
  D.cdc = pool.up(coarse->dep_conn_);
  D.cdw = pool.up(coarse->dep_w_);
  int *first = pool.get<int>(n), *len = pool.get<int>(n + 1), *rowp = pool.get<int>(n + 1);
  int *bad = pool.get<int>(1);
  if (!D.fo || !D.fconn || !D.co || !D.cconn || !D.cptr || !D.cdc || !D.cdw || !first || !len ||
      !rowp || !bad)
    return fail("device allocation failed");
  (void)hipMemset(first, 0x7f, (n ? n : 1) * sizeof(int));
  (void)hipMemset(len, 0, (n + 1) * sizeof(int));
  (void)hipMemset(bad, 0, sizeof(int));
  hipLaunchKernelGGL(k_first_visit, dim3(blocks(m)), dim3(256), 0, 0, m, D.fconn, first);
  // lengths in len[1..n], so that the inclusive scan is the row pointer
  if (n > 0)
    hipLaunchKernelGGL((k_interp_rows<false>), dim3(blocks(n)), dim3(256), 0, 0, D, n, first,
                       len + 1, nullptr, nullptr, nullptr, bad);
  size_t tb = 0;
  (void)hipcub::DeviceScan::InclusiveSum(nullptr, tb, len, rowp, n + 1);
  void *tmp = pool.get<unsigned char>(tb);
  if (!tmp) return fail("device allocation failed");
  if (hipcub::DeviceScan::InclusiveSum(tmp, tb, len, rowp, n + 1) != hipSuccess)
    return fail("scan failed");
  int hbad = 0, nnz = 0;
  if (hipMemcpy(&hbad, bad, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess ||
      hipMemcpy(&nnz, rowp + n, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess)
    return fail("copy failed");
  if (hbad) {
    last_error = "create_interpolation: no enclosing coarse octant (forests do not nest)";
    return -1;
  }
  int *cols = pool.get<int>(nnz);
  double *w = pool.get<double>(nnz);
  if (!cols || !w) return fail("device allocation failed");
  if (n > 0)
    hipLaunchKernelGGL((k_interp_rows<true>), dim3(blocks(n)), dim3(256), 0, 0, D, n, first,
                       nullptr, rowp, cols, w, bad);
  if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess)
    return fail("kernel launch failed");
  P.nrows = n;
  P.ncols = coarse->num_owned_;
  P.rowp.resize(n + 1);
  P.cols.resize(nnz);
  P.w.resize(nnz);
  if (hipMemcpy(P.rowp.data(), rowp, (n + 1) * sizeof(int), hipMemcpyDeviceToHost) != hipSuccess ||
      (nnz > 0 &&
       (hipMemcpy(P.cols.data(), cols, nnz * sizeof(int), hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(P.w.data(), w, nnz * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)))
    return fail("copy failed");
  if (ms) *ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return 0;
}

}  // namespace tmrx
