// Surface output of a nodal design field: TMR_GenerateBinFile and
// TMR_ConvertBinToSTL (src/TMR_STLTools.h:55-77).  The reference writes the
// level set {x = cutoff} of the filter forest's field by marching cubes plus
// the domain-boundary faces of the solid region.  Here each element is split
// into six tetrahedra around its 0-7 diagonal and the level set is traced
// with marching tetrahedra (exact for fields linear in an element's
// tetrahedra, no case table); boundary faces of elements on outer block
// faces are clipped to the solid region {x >= cutoff}.  Triangles are
// oriented with normals out of the solid.  File formats as the reference's:
//   binary: int32 triangle count, then 9 doubles per triangle (3 vertices);
//   ASCII STL: "solid topology", one facet per triangle with its unit normal.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/tmrx.h"
#include "forest.h"

namespace tmrx {
void set_error(const std::string &s);
}
tmrx::Forest *tmrx_forest_impl(tmrx_forest *f);

namespace {

struct V3 {
  double x, y, z;
};
V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
V3 lerp(V3 a, V3 b, double t) { return {a.x + t * (b.x - a.x), a.y + t * (b.y - a.y), a.z + t * (b.z - a.z)}; }
V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

struct Tris {
  std::vector<double> v;  // 9 per triangle
  // add (a, b, c), flipped if its normal points against `out`
  void add(V3 a, V3 b, V3 c, V3 out) {
    const V3 n = cross(sub(b, a), sub(c, a));
    if (dot(n, n) == 0.0) return;  // degenerate (a vertex on the level set)
    if (dot(n, out) < 0.0) std::swap(b, c);
    const double t[9] = {a.x, a.y, a.z, b.x, b.y, b.z, c.x, c.y, c.z};
    v.insert(v.end(), t, t + 9);
  }
};

// point on edge (p, q) where the field crosses the cutoff
V3 cut(V3 p, double vp, V3 q, double vq, double c) { return lerp(p, q, (c - vp) / (vq - vp)); }

// marching tetrahedra: the part of the level set inside tet (P, v)
void tet(const V3 P[4], const double v[4], double c, Tris &T) {
  int in[4], out[4], ni = 0, no = 0;
  for (int k = 0; k < 4; k++) {
    if (v[k] >= c)
      in[ni++] = k;
    else
      out[no++] = k;
  }
  if (ni == 0 || ni == 4) return;
  V3 ci = {0, 0, 0}, co = {0, 0, 0};
  for (int k = 0; k < ni; k++) ci = {ci.x + P[in[k]].x / ni, ci.y + P[in[k]].y / ni, ci.z + P[in[k]].z / ni};
  for (int k = 0; k < no; k++) co = {co.x + P[out[k]].x / no, co.y + P[out[k]].y / no, co.z + P[out[k]].z / no};
  const V3 dir = sub(co, ci);  // out of the solid
  if (ni == 1 || ni == 3) {
    const int s = ni == 1 ? in[0] : out[0];
    const int *o = ni == 1 ? out : in;
    V3 q[3];
    for (int k = 0; k < 3; k++) q[k] = cut(P[s], v[s], P[o[k]], v[o[k]], c);
    T.add(q[0], q[1], q[2], dir);
  } else {
    const V3 a = cut(P[in[0]], v[in[0]], P[out[0]], v[out[0]], c);
    const V3 b = cut(P[in[0]], v[in[0]], P[out[1]], v[out[1]], c);
    const V3 d = cut(P[in[1]], v[in[1]], P[out[1]], v[out[1]], c);
    const V3 e = cut(P[in[1]], v[in[1]], P[out[0]], v[out[0]], c);
    T.add(a, b, d, dir);
    T.add(a, d, e, dir);
  }
}

// the part of triangle (P, v) where v >= c, fan-triangulated, normal along out
void clip_tri(const V3 P[3], const double v[3], double c, V3 out, Tris &T) {
  V3 poly[4];
  int n = 0;
  for (int k = 0; k < 3; k++) {
    const int j = (k + 1) % 3;
    const bool ik = v[k] >= c, ij = v[j] >= c;
    if (ik) poly[n++] = P[k];
    if (ik != ij) poly[n++] = cut(P[k], v[k], P[j], v[j], c);
  }
  for (int k = 1; k + 1 < n; k++) T.add(poly[0], poly[k], poly[k + 1], out);
}

}  // namespace

extern "C" int tmrx_forest_generate_bin_file(tmrx_forest *filter, const char *filename,
                                             const double *x, int vars_per_node, int x_offset,
                                             double cutoff, int *ntris) {
  tmrx::Forest *f = tmrx_forest_impl(filter);
  if (!f || !filename || !x || vars_per_node < 1 || x_offset < 0 || x_offset >= vars_per_node) {
    tmrx::set_error("tmrx_forest_generate_bin_file: invalid arguments");
    return -1;
  }
  if (f->create_nodes()) {
    tmrx::set_error("tmrx_forest_generate_bin_file: " + f->last_error);
    return -1;
  }
  const auto &octs = f->octants();
  const auto &conn = f->conn();
  const auto &dp = f->dep_ptr();
  const auto &dc = f->dep_conn();
  const auto &dw = f->dep_weights();
  const auto &X = f->points();
  const int ndep = f->num_dep_nodes();
  const tmrx::BlockConn &bd = f->conn_data();
  auto pos = [&](int g) {
    const size_t i = g >= 0 ? (size_t)(g + ndep) : (size_t)(ndep - 1 - (-g - 1));
    return V3{X[3 * i], X[3 * i + 1], X[3 * i + 2]};
  };
  auto val = [&](int g) {
    if (g >= 0) return x[(size_t)vars_per_node * g + x_offset];
    double s = 0.0;
    for (int j = dp[-g - 1]; j < dp[-g]; j++) s += dw[j] * x[(size_t)vars_per_node * dc[j] + x_offset];
    return s;
  };
  // six tetrahedra around the 0-7 diagonal (corner bit d = +1 side of axis d)
  static const int TET[6][4] = {{0, 1, 3, 7}, {0, 3, 2, 7}, {0, 2, 6, 7},
                                {0, 6, 4, 7}, {0, 4, 5, 7}, {0, 5, 1, 7}};
  Tris T;
  for (size_t e = 0; e < octs.size(); e++) {
    V3 P[8];
    double v[8];
    for (int a = 0; a < 8; a++) {
      P[a] = pos(conn[8 * e + a]);
      v[a] = val(conn[8 * e + a]);
    }
    for (int t = 0; t < 6; t++) {
      V3 Q[4];
      double w[4];
      for (int k = 0; k < 4; k++) {
        Q[k] = P[TET[t][k]];
        w[k] = v[TET[t][k]];
      }
      tet(Q, w, cutoff, T);
    }
    // faces of the element on outer block faces (faces shared by no other block)
    const tmrx::Oct &o = octs[e];
    const int32_t h = 1 << (tmrx::MAX_LEVEL - o.level);
    const int32_t coord[3] = {o.x, o.y, o.z};
    V3 cen = {0, 0, 0};
    for (int a = 0; a < 8; a++) cen = {cen.x + P[a].x / 8, cen.y + P[a].y / 8, cen.z + P[a].z / 8};
    for (int fi = 0; fi < 6; fi++) {
      const int ax = fi / 2, hi = fi % 2;
      if (coord[ax] != (hi ? tmrx::HMAX - h : 0)) continue;
      const int uf = bd.block_face_conn[6 * o.block + fi];
      if (bd.face_block_ptr[uf + 1] - bd.face_block_ptr[uf] != 1) continue;
      int cor[4], n = 0;
      for (int a = 0; a < 8; a++)
        if (((a >> ax) & 1) == hi) cor[n++] = a;
      // corners in cyclic order: (0, 1, 3, 2) of the four on the face
      const int cyc[4] = {cor[0], cor[1], cor[3], cor[2]};
      V3 fc = {0, 0, 0};
      for (int k = 0; k < 4; k++) fc = {fc.x + P[cyc[k]].x / 4, fc.y + P[cyc[k]].y / 4, fc.z + P[cyc[k]].z / 4};
      const V3 outward = sub(fc, cen);
      for (int tri = 0; tri < 2; tri++) {
        const int i0 = cyc[0], i1 = cyc[1 + tri], i2 = cyc[2 + tri];
        const V3 Q[3] = {P[i0], P[i1], P[i2]};
        const double w[3] = {v[i0], v[i1], v[i2]};
        clip_tri(Q, w, cutoff, outward, T);
      }
    }
  }
  const int nt = (int)(T.v.size() / 9);
  FILE *fp = std::fopen(filename, "wb");
  if (!fp) {
    tmrx::set_error(std::string("tmrx_forest_generate_bin_file: cannot open ") + filename);
    return -1;
  }
  bool ok = std::fwrite(&nt, sizeof(int), 1, fp) == 1;
  if (nt) ok = ok && std::fwrite(T.v.data(), sizeof(double), T.v.size(), fp) == T.v.size();
  ok = (std::fclose(fp) == 0) && ok;
  if (!ok) {
    tmrx::set_error("tmrx_forest_generate_bin_file: write failed");
    return -1;
  }
  if (ntris) *ntris = nt;
  return 0;
}

extern "C" int tmrx_convert_bin_to_stl(const char *binfile, const char *stlfile) {
  if (!binfile || !stlfile) {
    tmrx::set_error("tmrx_convert_bin_to_stl: null file name");
    return -1;
  }
  FILE *fp = std::fopen(binfile, "rb");
  if (!fp) {
    tmrx::set_error(std::string("tmrx_convert_bin_to_stl: cannot open ") + binfile);
    return -1;
  }
  int nt = 0;
  std::vector<double> v;
  bool ok = std::fread(&nt, sizeof(int), 1, fp) == 1 && nt >= 0;
  if (ok) {
    v.resize(9 * (size_t)nt);
    ok = std::fread(v.data(), sizeof(double), v.size(), fp) == v.size();
  }
  std::fclose(fp);
  if (!ok) {
    tmrx::set_error("tmrx_convert_bin_to_stl: truncated or invalid file");
    return -1;
  }
  FILE *out = std::fopen(stlfile, "w");
  if (!out) {
    tmrx::set_error(std::string("tmrx_convert_bin_to_stl: cannot open ") + stlfile);
    return -1;
  }
  std::fprintf(out, "solid topology\n");
  for (int i = 0; i < nt; i++) {
    const double *t = &v[9 * (size_t)i];
    const V3 a{t[0], t[1], t[2]}, b{t[3], t[4], t[5]}, c{t[6], t[7], t[8]};
    V3 n = cross(sub(b, a), sub(c, a));
    const double l = std::sqrt(dot(n, n));
    if (l > 0) n = {n.x / l, n.y / l, n.z / l};
    std::fprintf(out, "facet normal %e %e %e\nouter loop\n", n.x, n.y, n.z);
    for (int k = 0; k < 3; k++) std::fprintf(out, "vertex %e %e %e\n", t[3 * k], t[3 * k + 1], t[3 * k + 2]);
    std::fprintf(out, "endloop\nendfacet\n");
  }
  std::fprintf(out, "endsolid topology\n");
  if (std::fclose(out) != 0) {
    tmrx::set_error("tmrx_convert_bin_to_stl: write failed");
    return -1;
  }
  return 0;
}

// Volume output in place of TACS's ToFH5 (examples/topology_optimization/
// cantilever/cantilever.py:154-182 writes connectivity, nodes and element
// extras to an .f5 file; TACS and its format are not part of the
// reference): a legacy-VTK unstructured grid of the forest's hex8 elements
// with an optional nodal field (node_vpn values per independent node,
// dependent nodes interpolated from their parents) and an optional element
// field (elem_vpe values per element).  Points follow getPoints's order
// (dependent nodes first).  VTK's hexahedron corner order is 0 1 3 2 4 5 7 6
// in TMR's local numbering (x fastest).
extern "C" int tmrx_forest_write_vtk(tmrx_forest *forest, const char *filename,
                                     const double *node_vals, int node_vpn, const char *node_name,
                                     const double *elem_vals, int elem_vpe,
                                     const char *elem_name) {
  tmrx::Forest *f = tmrx_forest_impl(forest);
  if (!f || !filename || (node_vals && node_vpn < 1) || (elem_vals && elem_vpe < 1)) {
    tmrx::set_error("tmrx_forest_write_vtk: invalid arguments");
    return -1;
  }
  if (f->create_nodes()) {
    tmrx::set_error("tmrx_forest_write_vtk: " + f->last_error);
    return -1;
  }
  const auto &conn = f->conn();
  const auto &dp = f->dep_ptr();
  const auto &dc = f->dep_conn();
  const auto &dw = f->dep_weights();
  const auto &X = f->points();
  const int ndep = f->num_dep_nodes();
  const int nloc = (int)f->node_numbers().size();
  const int ne = f->num_elements();
  auto index = [&](int g) { return g >= 0 ? g + ndep : ndep - 1 - (-g - 1); };
  FILE *fp = std::fopen(filename, "w");
  if (!fp) {
    tmrx::set_error(std::string("tmrx_forest_write_vtk: cannot open ") + filename);
    return -1;
  }
  std::fprintf(fp, "# vtk DataFile Version 3.0\ntmrx forest\nASCII\nDATASET UNSTRUCTURED_GRID\n");
  std::fprintf(fp, "POINTS %d double\n", nloc);
  for (int i = 0; i < nloc; i++)
    std::fprintf(fp, "%.17g %.17g %.17g\n", X[3 * (size_t)i], X[3 * (size_t)i + 1], X[3 * (size_t)i + 2]);
  std::fprintf(fp, "CELLS %d %d\n", ne, 9 * ne);
  static const int VTK_ORDER[8] = {0, 1, 3, 2, 4, 5, 7, 6};
  for (int e = 0; e < ne; e++) {
    std::fprintf(fp, "8");
    for (int a = 0; a < 8; a++) std::fprintf(fp, " %d", index(conn[8 * (size_t)e + VTK_ORDER[a]]));
    std::fprintf(fp, "\n");
  }
  std::fprintf(fp, "CELL_TYPES %d\n", ne);
  for (int e = 0; e < ne; e++) std::fprintf(fp, "12\n");
  if (node_vals) {
    std::fprintf(fp, "POINT_DATA %d\n", nloc);
    if (node_vpn == 3)
      std::fprintf(fp, "VECTORS %s double\n", node_name ? node_name : "u");
    else
      std::fprintf(fp, "SCALARS %s double %d\nLOOKUP_TABLE default\n", node_name ? node_name : "x",
                   node_vpn);
    std::vector<double> v(node_vpn);
    for (int i = 0; i < nloc; i++) {
      const int g = i >= ndep ? i - ndep : -(ndep - i);  // the node number at point i
      std::fill(v.begin(), v.end(), 0.0);
      if (g >= 0) {
        for (int k = 0; k < node_vpn; k++) v[k] = node_vals[(size_t)node_vpn * g + k];
      } else {
        for (int j = dp[-g - 1]; j < dp[-g]; j++)
          for (int k = 0; k < node_vpn; k++) v[k] += dw[j] * node_vals[(size_t)node_vpn * dc[j] + k];
      }
      for (int k = 0; k < node_vpn; k++) std::fprintf(fp, k ? " %.17g" : "%.17g", v[k]);
      std::fprintf(fp, "\n");
    }
  }
  if (elem_vals) {
    std::fprintf(fp, "CELL_DATA %d\nSCALARS %s double %d\nLOOKUP_TABLE default\n", ne,
                 elem_name ? elem_name : "rho", elem_vpe);
    for (int e = 0; e < ne; e++) {
      for (int k = 0; k < elem_vpe; k++)
        std::fprintf(fp, k ? " %.17g" : "%.17g", elem_vals[(size_t)elem_vpe * e + k]);
      std::fprintf(fp, "\n");
    }
  }
  if (std::fclose(fp) != 0) {
    tmrx::set_error("tmrx_forest_write_vtk: write failed");
    return -1;
  }
  return 0;
}
