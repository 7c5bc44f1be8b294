/*
  Geometric CPU oracle of the forest index maps -- TEST INFRASTRUCTURE ONLY.

  Only tests/ and __graft_entry__.smoke() load this; the product never does.

  Scope: forests whose blocks form an nx x ny x nz lattice of equally oriented
  boxes (block b = i + nx (j + ny k), corner c = i + 2j + 4k), which is every
  forest the workloads use.  Each function restates the *result* of a
  reference routine by brute force over global lattice coordinates instead of
  repeating the reference's local, rank-aware algorithm, so that the two
  implementations share only the definitions:

    octant order     TMROctant::compare/compareNode, src/TMROctant.cpp:171-274
                     (restated literally: most significant differing bit,
                     x before y before z, ties by level / info)
    sort+dedupe      TMROctantArray::sort, src/TMROctant.cpp:357-399 (keep the
                     last = finest octant of equal position)
    balance          TMROctForest::balance, src/TMROctForest.cpp:2917-3089:
                     the coarsest complete refinement, containing the input
                     families, in which leaves touching across a face or edge
                     (and a corner when balance_corner) differ by <= 1 level;
                     computed by repeated splitting of violating leaves
    nodes            TMROctForest::createNodes, src/TMROctForest.cpp:4064-4268:
                     element corners canonicalised to the lowest-index block
                     containing them (transformNode, :3847-4062, including the
                     2^30 -> 2^30 - 1 truncation); a node is dependent when it
                     lies strictly inside an edge or face of a coarser leaf;
                     dependents numbered -1, -2, ... and the rest 0, 1, ... in
                     node order (:4114-4120, 4174-4181)
    dep weights      createDependentConn, src/TMROctForest.cpp:5157-5519:
                     edge-hanging -> the 2 edge ends (lower first) x 1/2,
                     face-hanging -> the 4 face corners (u fastest) x 1/4
    interpolation    createInterpolation, src/TMROctForest.cpp:6611-6793: row
                     of fine independent node i = coarse basis functions
                     (trilinear on a coarse leaf containing the node, hanging
                     coarse nodes expanded) evaluated at the node, non-zero
                     entries, merged and sorted by column (TMRIndexWeight::
                     uniqueSort, src/TMRBase.h:112-135)
  Pins: tests/test_oracle_forest.py (closed-form node counts, 2:1 balance
  checked leaf by leaf, minimality, Morton order against the reference's own
  TMROctantArray::sort built in oracle/_ref, partition of unity and exact
  reproduction of trilinear fields by the interpolation).
*/
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define MAXL 30
#define HMAX (1 << MAXL)

typedef struct {
  int32_t block, x, y, z, tag;
  int16_t level, info;
} orc_oct;

/* ---- literal restatement of TMROctant::compare / compareNode ---- */
static int cmp_common(const orc_oct *a, const orc_oct *b, int *tie) {
  if (a->block != b->block) {
    *tie = 0;
    return a->block - b->block;
  }
  uint32_t xx = (uint32_t)(a->x ^ b->x), yy = (uint32_t)(a->y ^ b->y), zz = (uint32_t)(a->z ^ b->z);
  uint32_t sor = xx | yy | zz;
  if (sor == 0) {
    *tie = 1;
    return 0;
  }
  *tie = 0;
  int d;
  if (xx > (sor ^ xx))
    d = a->x - b->x;
  else if (yy > (sor ^ yy))
    d = a->y - b->y;
  else
    d = a->z - b->z;
  return d > 0 ? 1 : (d < 0 ? -1 : 0);
}
static int cmp_oct(const void *pa, const void *pb) {
  int tie;
  const orc_oct *a = (const orc_oct *)pa, *b = (const orc_oct *)pb;
  int c = cmp_common(a, b, &tie);
  return tie ? a->level - b->level : c;
}
static int cmp_pos(const orc_oct *a, const orc_oct *b) {
  int tie;
  int c = cmp_common(a, b, &tie);
  return tie ? 0 : c;
}
static int cmp_node(const void *pa, const void *pb) {
  int tie;
  const orc_oct *a = (const orc_oct *)pa, *b = (const orc_oct *)pb;
  int c = cmp_common(a, b, &tie);
  return tie ? a->info - b->info : c;
}

int orc_sort_octants(orc_oct *a, int n, int use_node) {
  qsort(a, (size_t)n, sizeof(orc_oct), use_node ? cmp_node : cmp_oct);
  int j = 0;
  for (int i = 0; i < n; i++, j++) {
    while (i < n - 1 && (use_node ? cmp_node(&a[i], &a[i + 1]) == 0 : cmp_pos(&a[i], &a[i + 1]) == 0)) i++;
    if (i != j) a[j] = a[i];
  }
  return j;
}

/* ---- hash set of octants (block, x, y, z, level) ---- */
typedef struct {
  int64_t *keys; /* packed; -1 empty */
  size_t cap, n;
} hset;

static int64_t okey(int b, int32_t x, int32_t y, int32_t z, int l) {
  /* the oracle is used with levels <= 12: coordinates are multiples of
     2^(30 - 12) and fit in 12 bits each after the shift */
  const int s = MAXL - 12;
  return ((((((int64_t)b * 16 + l) << 12 | (int64_t)(x >> s)) << 12) | (int64_t)(y >> s)) << 12) |
         (int64_t)(z >> s);
}
static uint64_t mix(uint64_t k) {
  k ^= k >> 33;
  k *= 0xff51afd7ed558ccdULL;
  k ^= k >> 33;
  k *= 0xc4ceb9fe1a85ec53ULL;
  k ^= k >> 33;
  return k;
}
static void hs_init(hset *h, size_t cap) {
  h->cap = 1;
  while (h->cap < 2 * cap + 16) h->cap <<= 1;
  h->keys = (int64_t *)malloc(sizeof(int64_t) * h->cap);
  memset(h->keys, 0xff, sizeof(int64_t) * h->cap);
  h->n = 0;
}
static int hs_has(const hset *h, int64_t k) {
  size_t i = mix((uint64_t)k) & (h->cap - 1);
  while (h->keys[i] != -1) {
    if (h->keys[i] == k) return 1;
    i = (i + 1) & (h->cap - 1);
  }
  return 0;
}
static void hs_grow(hset *h);
static int hs_add(hset *h, int64_t k) {
  if (2 * (h->n + 1) > h->cap) hs_grow(h);
  size_t i = mix((uint64_t)k) & (h->cap - 1);
  while (h->keys[i] != -1) {
    if (h->keys[i] == k) return 0;
    i = (i + 1) & (h->cap - 1);
  }
  h->keys[i] = k;
  h->n++;
  return 1;
}
static void hs_grow(hset *h) {
  int64_t *old = h->keys;
  size_t oc = h->cap;
  h->cap *= 2;
  h->keys = (int64_t *)malloc(sizeof(int64_t) * h->cap);
  memset(h->keys, 0xff, sizeof(int64_t) * h->cap);
  h->n = 0;
  for (size_t i = 0; i < oc; i++)
    if (old[i] != -1) hs_add(h, old[i]);
  free(old);
}

/* ---- lattice geometry ---- */
static int g_nx, g_ny, g_nz;
static hset g_ref; /* refined octants */

static void blk_ijk(int b, int *i, int *j, int *k) {
  *i = b % g_nx;
  *j = (b / g_nx) % g_ny;
  *k = b / (g_nx * g_ny);
}

/* leaf containing the global point (gx, gy, gz) (int64 in units of 2^-30
   block), following refinement from the root; returns 0 if outside */
static int locate(int64_t gx, int64_t gy, int64_t gz, orc_oct *leaf) {
  if (gx < 0 || gy < 0 || gz < 0) return 0;
  int i = (int)(gx >> MAXL), j = (int)(gy >> MAXL), k = (int)(gz >> MAXL);
  if (i >= g_nx || j >= g_ny || k >= g_nz) return 0;
  int b = i + g_nx * (j + g_ny * k);
  int32_t x = (int32_t)(gx - ((int64_t)i << MAXL)), y = (int32_t)(gy - ((int64_t)j << MAXL)),
          z = (int32_t)(gz - ((int64_t)k << MAXL));
  int l = 0;
  int32_t ox = 0, oy = 0, oz = 0;
  while (hs_has(&g_ref, okey(b, ox, oy, oz, l))) {
    l++;
    int32_t h = 1 << (MAXL - l);
    ox = x & ~(h - 1);
    oy = y & ~(h - 1);
    oz = z & ~(h - 1);
  }
  leaf->block = b;
  leaf->x = ox;
  leaf->y = oy;
  leaf->z = oz;
  leaf->level = (int16_t)l;
  leaf->tag = 0;
  leaf->info = 0;
  return 1;
}

static void mark_refined_with_ancestors(int b, int32_t x, int32_t y, int32_t z, int l) {
  while (l >= 0) {
    int32_t h = 1 << (MAXL - l);
    x &= ~(h - 1);
    y &= ~(h - 1);
    z &= ~(h - 1);
    if (!hs_add(&g_ref, okey(b, x, y, z, l))) return;
    l--;
  }
}

static orc_oct *collect_leaves(int *nout) {
  /* DFS from every root */
  size_t cap = 1024, n = 0;
  orc_oct *out = (orc_oct *)malloc(sizeof(orc_oct) * cap);
  orc_oct *stack = (orc_oct *)malloc(sizeof(orc_oct) * 8 * (MAXL + 2) * 8);
  int nb = g_nx * g_ny * g_nz;
  for (int b = 0; b < nb; b++) {
    int sp = 0;
    orc_oct r = {b, 0, 0, 0, 0, 0, 0};
    stack[sp++] = r;
    while (sp) {
      orc_oct o = stack[--sp];
      if (hs_has(&g_ref, okey(o.block, o.x, o.y, o.z, o.level))) {
        int32_t h = 1 << (MAXL - o.level - 1);
        for (int c = 7; c >= 0; c--) {
          orc_oct ch = o;
          ch.level = o.level + 1;
          ch.x = o.x + ((c & 1) ? h : 0);
          ch.y = o.y + ((c & 2) ? h : 0);
          ch.z = o.z + ((c & 4) ? h : 0);
          stack[sp++] = ch;
        }
      } else {
        if (n == cap) {
          cap *= 2;
          out = (orc_oct *)realloc(out, sizeof(orc_oct) * cap);
        }
        out[n++] = o;
      }
    }
  }
  free(stack);
  *nout = (int)n;
  return out;
}

static orc_oct *g_leaves = NULL;
static int g_nleaves = 0;

/* Balance by repeated splitting.  in: octants whose families must exist
   (the output of createTrees/refine/coarsen).  Returns the leaf count; the
   sorted leaves are fetched with orc_get_leaves. */
int orc_balance(int nx, int ny, int nz, const orc_oct *in, int n, int corner) {
  g_nx = nx;
  g_ny = ny;
  g_nz = nz;
  if (g_ref.keys) free(g_ref.keys);
  hs_init(&g_ref, (size_t)n + 16);
  for (int i = 0; i < n; i++)
    if (in[i].level > 0) {
      int32_t h = 1 << (MAXL - in[i].level + 1);
      mark_refined_with_ancestors(in[i].block, in[i].x & ~(h - 1), in[i].y & ~(h - 1),
                                  in[i].z & ~(h - 1), in[i].level - 1);
    }
  for (;;) {
    int nl;
    orc_oct *L = collect_leaves(&nl);
    int changed = 0;
    for (int i = 0; i < nl; i++) {
      const orc_oct *q = &L[i];
      int bi, bj, bk;
      blk_ijk(q->block, &bi, &bj, &bk);
      int64_t h = (int64_t)1 << (MAXL - q->level);
      int64_t gx = ((int64_t)bi << MAXL) + q->x, gy = ((int64_t)bj << MAXL) + q->y,
              gz = ((int64_t)bk << MAXL) + q->z;
      for (int dz = -1; dz <= 1; dz++)
        for (int dy = -1; dy <= 1; dy++)
          for (int dx = -1; dx <= 1; dx++) {
            int nz0 = (dx != 0) + (dy != 0) + (dz != 0);
            if (nz0 == 0 || (nz0 == 3 && !corner)) continue;
            /* a point just across the face/edge/corner, at the centre of the
               shared feature */
            int64_t px = dx < 0 ? gx - 1 : dx > 0 ? gx + h : gx + h / 2;
            int64_t py = dy < 0 ? gy - 1 : dy > 0 ? gy + h : gy + h / 2;
            int64_t pz = dz < 0 ? gz - 1 : dz > 0 ? gz + h : gz + h / 2;
            orc_oct nb;
            if (!locate(px, py, pz, &nb)) continue;
            if (nb.level < q->level - 1) {
              mark_refined_with_ancestors(nb.block, nb.x, nb.y, nb.z, nb.level);
              changed = 1;
            }
          }
    }
    if (!changed) {
      if (g_leaves) free(g_leaves);
      g_leaves = L;
      g_nleaves = orc_sort_octants(L, nl, 0);
      for (int i = 0; i < g_nleaves; i++) g_leaves[i].tag = i;
      return g_nleaves;
    }
    free(L);
  }
}

int orc_get_leaves(orc_oct *out) {
  memcpy(out, g_leaves, sizeof(orc_oct) * (size_t)g_nleaves);
  return g_nleaves;
}

/* is octant o (global) refined in the current refinement set */
int orc_is_refined(const orc_oct *o) { return hs_has(&g_ref, okey(o->block, o->x, o->y, o->z, o->level)); }

/* ---- nodes ---- */
typedef struct {
  orc_oct o;     /* canonical node octant (block + local coords) */
  int64_t g[3];  /* global coordinates */
} gnode;

static int cmp_gnode(const void *a, const void *b) {
  return cmp_node(&((const gnode *)a)->o, &((const gnode *)b)->o);
}

static void canon(int64_t gx, int64_t gy, int64_t gz, orc_oct *o) {
  /* lowest-index block whose closed box contains the point */
  int best = -1;
  int64_t G[3] = {gx, gy, gz};
  int dims[3] = {g_nx, g_ny, g_nz};
  int lo[3], hi[3];
  for (int a = 0; a < 3; a++) {
    int64_t q = G[a] >> MAXL, r = G[a] & (HMAX - 1);
    hi[a] = (int)q < dims[a] ? (int)q : dims[a] - 1;
    lo[a] = (r == 0 && q > 0) ? (int)q - 1 : hi[a];
  }
  for (int k = lo[2]; k <= hi[2]; k++)
    for (int j = lo[1]; j <= hi[1]; j++)
      for (int i = lo[0]; i <= hi[0]; i++) {
        int b = i + g_nx * (j + g_ny * k);
        if (best < 0 || b < best) best = b;
      }
  int bi, bj, bk;
  blk_ijk(best, &bi, &bj, &bk);
  int64_t x = gx - ((int64_t)bi << MAXL), y = gy - ((int64_t)bj << MAXL), z = gz - ((int64_t)bk << MAXL);
  o->block = best;
  o->x = (int32_t)(x == HMAX ? HMAX - 1 : x);
  o->y = (int32_t)(y == HMAX ? HMAX - 1 : y);
  o->z = (int32_t)(z == HMAX ? HMAX - 1 : z);
  o->tag = 0;
  o->level = 0;
  o->info = 0;
}

static gnode *g_nodes = NULL;
static int g_nnodes = 0, g_nindep = 0, g_ndep = 0;
static int *g_num = NULL;     /* node index -> number */
static int *g_conn = NULL;    /* 8 per leaf, numbers */
static int *g_dptr = NULL, *g_dconn = NULL;
static double *g_dw = NULL;

static int find_gnode(int64_t gx, int64_t gy, int64_t gz) {
  gnode key;
  canon(gx, gy, gz, &key.o);
  gnode *f = (gnode *)bsearch(&key, g_nodes, (size_t)g_nnodes, sizeof(gnode), cmp_gnode);
  return f ? (int)(f - g_nodes) : -1;
}

static void leaf_global(const orc_oct *q, int64_t g[3], int64_t *h) {
  int bi, bj, bk;
  blk_ijk(q->block, &bi, &bj, &bk);
  g[0] = ((int64_t)bi << MAXL) + q->x;
  g[1] = ((int64_t)bj << MAXL) + q->y;
  g[2] = ((int64_t)bk << MAXL) + q->z;
  *h = (int64_t)1 << (MAXL - q->level);
}

/* hanging test: returns number of parents (0, 2, 4) and their global points */
static int hanging(const int64_t P[3], int64_t par[4][3]) {
  for (int c = 0; c < 8; c++) {
    orc_oct q;
    int64_t p0 = P[0] - ((c & 1) ? 1 : 0), p1 = P[1] - ((c & 2) ? 1 : 0), p2 = P[2] - ((c & 4) ? 1 : 0);
    if (!locate(p0, p1, p2, &q)) continue;
    int64_t g[3], h;
    leaf_global(&q, g, &h);
    int onb = 0, inside = 0, axis_in[3];
    for (int a = 0; a < 3; a++) {
      int at_end = (P[a] == g[a] || P[a] == g[a] + h);
      axis_in[a] = !at_end;
      if (at_end) onb++;
      else inside++;
    }
    if (onb == 3) continue;  /* corner of q */
    if (onb == 2) {          /* inside an edge of q along the free axis */
      int a = axis_in[0] ? 0 : axis_in[1] ? 1 : 2;
      for (int e = 0; e < 2; e++) {
        for (int d = 0; d < 3; d++) par[e][d] = P[d];
        par[e][a] = g[a] + e * h;
      }
      return 2;
    }
    if (onb == 1) { /* inside a face of q */
      int fa = !axis_in[0] ? 0 : !axis_in[1] ? 1 : 2;
      int ua = (fa == 0) ? 1 : 0, va = (fa == 2) ? 1 : 2;
      for (int vv = 0; vv < 2; vv++)
        for (int uu = 0; uu < 2; uu++) {
          int e = uu + 2 * vv;
          par[e][fa] = P[fa];
          par[e][ua] = g[ua] + uu * h;
          par[e][va] = g[va] + vv * h;
        }
      return 4;
    }
    (void)inside;
  }
  return 0;
}

/* Nodes of the current leaves (after orc_balance).  Returns the number of
   independent nodes; sizes through orc_node_sizes. */
int orc_create_nodes(void) {
  free(g_nodes);
  free(g_num);
  free(g_conn);
  free(g_dptr);
  free(g_dconn);
  free(g_dw);
  int n = g_nleaves;
  gnode *all = (gnode *)malloc(sizeof(gnode) * 8 * (size_t)n);
  for (int i = 0; i < n; i++) {
    int64_t g[3], h;
    leaf_global(&g_leaves[i], g, &h);
    for (int c = 0; c < 8; c++) {
      gnode *t = &all[8 * i + c];
      t->g[0] = g[0] + ((c & 1) ? h : 0);
      t->g[1] = g[1] + ((c & 2) ? h : 0);
      t->g[2] = g[2] + ((c & 4) ? h : 0);
      canon(t->g[0], t->g[1], t->g[2], &t->o);
    }
  }
  qsort(all, 8 * (size_t)n, sizeof(gnode), cmp_gnode);
  int m = 0;
  for (int i = 0; i < 8 * n; i++) {
    if (m > 0 && cmp_gnode(&all[i], &all[m - 1]) == 0) continue;
    all[m++] = all[i];
  }
  g_nodes = all;
  g_nnodes = m;
  g_num = (int *)malloc(sizeof(int) * (size_t)m);
  int64_t (*par)[4][3] = malloc(sizeof(*par) * (size_t)m);
  int *npar = (int *)malloc(sizeof(int) * (size_t)m);
  g_ndep = 0;
  g_nindep = 0;
  for (int i = 0; i < m; i++) {
    npar[i] = hanging(g_nodes[i].g, par[i]);
    if (npar[i]) g_num[i] = -(++g_ndep);
  }
  for (int i = 0; i < m; i++)
    if (!npar[i]) g_num[i] = g_nindep++;
  g_conn = (int *)malloc(sizeof(int) * 8 * (size_t)n);
  for (int i = 0; i < n; i++) {
    int64_t g[3], h;
    leaf_global(&g_leaves[i], g, &h);
    for (int c = 0; c < 8; c++) {
      int idx = find_gnode(g[0] + ((c & 1) ? h : 0), g[1] + ((c & 2) ? h : 0), g[2] + ((c & 4) ? h : 0));
      g_conn[8 * i + c] = g_num[idx];
    }
  }
  g_dptr = (int *)calloc((size_t)g_ndep + 1, sizeof(int));
  for (int i = 0; i < m; i++)
    if (npar[i]) g_dptr[-g_num[i]] = npar[i];
  for (int d = 0; d < g_ndep; d++) g_dptr[d + 1] += g_dptr[d];
  g_dconn = (int *)malloc(sizeof(int) * ((size_t)g_dptr[g_ndep] + 1));
  g_dw = (double *)malloc(sizeof(double) * ((size_t)g_dptr[g_ndep] + 1));
  for (int i = 0; i < m; i++) {
    if (!npar[i]) continue;
    int d = -g_num[i] - 1;
    for (int k = 0; k < npar[i]; k++) {
      int idx = find_gnode(par[i][k][0], par[i][k][1], par[i][k][2]);
      g_dconn[g_dptr[d] + k] = idx >= 0 ? g_num[idx] : -999999;
      g_dw[g_dptr[d] + k] = 1.0 / npar[i];
    }
  }
  free(par);
  free(npar);
  return g_nindep;
}

void orc_node_sizes(int *nnodes, int *nindep, int *ndep, int *ndepconn) {
  *nnodes = g_nnodes;
  *nindep = g_nindep;
  *ndep = g_ndep;
  *ndepconn = g_dptr ? g_dptr[g_ndep] : 0;
}

void orc_get_nodes(int *conn, int *dep_ptr, int *dep_conn, double *dep_w, int64_t *gxyz) {
  memcpy(conn, g_conn, sizeof(int) * 8 * (size_t)g_nleaves);
  memcpy(dep_ptr, g_dptr, sizeof(int) * ((size_t)g_ndep + 1));
  memcpy(dep_conn, g_dconn, sizeof(int) * (size_t)g_dptr[g_ndep]);
  memcpy(dep_w, g_dw, sizeof(double) * (size_t)g_dptr[g_ndep]);
  if (gxyz)
    for (int i = 0; i < g_nnodes; i++)
      for (int a = 0; a < 3; a++) gxyz[3 * i + a] = g_nodes[i].g[a];
}

/* numbers of the nodes lying on the closed outer face (axis a, side s) of the
   lattice box: getNodesWithName for a name given to those block faces */
int orc_nodes_on_box_face(int axis, int side, int *out) {
  int dims[3] = {g_nx, g_ny, g_nz};
  int64_t v = side ? ((int64_t)dims[axis] << MAXL) : 0;
  int n = 0;
  for (int i = 0; i < g_nnodes; i++)
    if (g_nodes[i].g[axis] == v) out[n++] = g_num[i];
  /* sorted ascending, unique */
  for (int i = 1; i < n; i++) {
    int t = out[i], j = i - 1;
    while (j >= 0 && out[j] > t) {
      out[j + 1] = out[j];
      j--;
    }
    out[j + 1] = t;
  }
  return n;
}

/* ---- interpolation between two oracle meshes ---- */
typedef struct {
  int nleaves, nnodes, nindep, ndep;
  orc_oct *leaves;
  gnode *nodes;
  int *num, *conn, *dptr, *dconn;
  double *dw;
  hset ref;
} mesh_t;

static mesh_t g_saved[2];

/* stash the current mesh in slot s (0 fine, 1 coarse) */
void orc_save_mesh(int s) {
  mesh_t *M = &g_saved[s];
  M->nleaves = g_nleaves;
  M->nnodes = g_nnodes;
  M->nindep = g_nindep;
  M->ndep = g_ndep;
  M->leaves = g_leaves;
  M->nodes = g_nodes;
  M->num = g_num;
  M->conn = g_conn;
  M->dptr = g_dptr;
  M->dconn = g_dconn;
  M->dw = g_dw;
  M->ref = g_ref;
  g_leaves = NULL;
  g_nodes = NULL;
  g_num = g_conn = g_dptr = g_dconn = NULL;
  g_dw = NULL;
  g_ref.keys = NULL;
  g_nleaves = g_nnodes = 0;
}

static int *g_prow = NULL, *g_pcol = NULL;
static double *g_pw = NULL;

typedef struct {
  int c;
  double w;
} cw_t;
static int cmp_cw(const void *a, const void *b) {
  int x = ((const cw_t *)a)->c, y = ((const cw_t *)b)->c;
  return x < y ? -1 : x > y;
}

/* P from saved coarse (slot 1) to saved fine (slot 0); returns nnz */
int orc_interp(void) {
  mesh_t *F = &g_saved[0], *Cm = &g_saved[1];
  free(g_prow);
  free(g_pcol);
  free(g_pw);
  g_prow = (int *)calloc((size_t)F->nindep + 1, sizeof(int));
  size_t cap = 64 * (size_t)F->nindep + 64, nnz = 0;
  g_pcol = (int *)malloc(sizeof(int) * cap);
  g_pw = (double *)malloc(sizeof(double) * cap);
  /* independent fine nodes in number order */
  int *order = (int *)malloc(sizeof(int) * (size_t)F->nindep);
  for (int i = 0; i < F->nnodes; i++)
    if (F->num[i] >= 0) order[F->num[i]] = i;
  hset save_ref = g_ref;
  g_ref = Cm->ref;
  for (int r = 0; r < F->nindep; r++) {
    const int64_t *P = F->nodes[order[r]].g;
    orc_oct q;
    int found = 0;
    for (int c = 0; c < 8 && !found; c++) {
      int64_t p0 = P[0] - ((c & 1) ? 1 : 0), p1 = P[1] - ((c & 2) ? 1 : 0), p2 = P[2] - ((c & 4) ? 1 : 0);
      found = locate(p0, p1, p2, &q);
    }
    int64_t g[3], h;
    leaf_global(&q, g, &h);
    /* coarse leaf index -> its conn (binary search in the sorted leaves) */
    orc_oct *lf = (orc_oct *)bsearch(&q, Cm->leaves, (size_t)Cm->nleaves, sizeof(orc_oct), cmp_oct);
    const int *cc = &Cm->conn[8 * (lf - Cm->leaves)];
    cw_t buf[64];
    int nb = 0;
    for (int c = 0; c < 8; c++) {
      double w = 1.0;
      for (int a = 0; a < 3; a++) {
        double t = (double)(P[a] - g[a]) / (double)h;
        w *= ((c >> a) & 1) ? t : 1.0 - t;
      }
      if (w == 0.0) continue;
      int n = cc[c];
      if (n >= 0) {
        buf[nb].c = n;
        buf[nb++].w = w;
      } else {
        int d = -n - 1;
        for (int k = Cm->dptr[d]; k < Cm->dptr[d + 1]; k++) {
          buf[nb].c = Cm->dconn[k];
          buf[nb++].w = w * Cm->dw[k];
        }
      }
    }
    qsort(buf, (size_t)nb, sizeof(cw_t), cmp_cw);
    int start = (int)nnz;
    for (int k = 0; k < nb; k++) {
      if ((int)nnz > start && g_pcol[nnz - 1] == buf[k].c) {
        g_pw[nnz - 1] += buf[k].w;
      } else {
        g_pcol[nnz] = buf[k].c;
        g_pw[nnz++] = buf[k].w;
      }
    }
    g_prow[r + 1] = (int)nnz;
  }
  g_ref = save_ref;
  free(order);
  return (int)nnz;
}

void orc_get_interp(int *rowp, int *cols, double *w) {
  mesh_t *F = &g_saved[0];
  memcpy(rowp, g_prow, sizeof(int) * ((size_t)F->nindep + 1));
  memcpy(cols, g_pcol, sizeof(int) * (size_t)g_prow[F->nindep]);
  memcpy(w, g_pw, sizeof(double) * (size_t)g_prow[F->nindep]);
}
