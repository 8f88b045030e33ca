// Mesh preprocessing: connectivity in the reference's numbering, geometry, and the level
// schedule of the element-order limiter sweep.
//
// Numbering rules (src/initialization.rs:112-261), which the device layout keeps so that
// every index the C ABI reports is the reference's:
//   * boundary edges come first, block by block, in file order; their node order is the file's;
//   * interior edges follow, numbered at first visit while scanning elements in order and each
//     element's local edges (n0,n1), (n1,n2), (n2,n0); an interior edge keeps the node order of the
//     element that created it;
//   * "left" (slot 0) of an interior edge is the element whose local edge runs ivertices[0] ->
//     ivertices[1]; a boundary edge's only element sits in slot 0;
//   * elements with a boundary edge are "boundary elements", the rest "internal".
// Geometry follows src/mesh.rs:15-84 (edge length and normal (dy, -dx)/|e| from ivertices[0] to
// ivertices[1]; J = (x2-x1)(y3-y1) - (x3-x1)(y2-y1); metric; circumradius; smallest altitude),
// and neighbours src/mesh.rs:133-147.
#include <algorithm>
#include <cmath>
#include <unordered_map>

#include "dgm_internal.h"

namespace dgm {

namespace {

struct PairHash {
  size_t operator()(uint64_t k) const { return std::hash<uint64_t>()(k * 0x9E3779B97F4A7C15ull); }
};

}  // namespace

int build_mesh(const dgm_mesh& in, HostMesh& m, std::string& err) {
  if (in.n_vertices <= 0 || !in.vertices_xy || in.n_element_blocks <= 0 || !in.element_block_count ||
      !in.element_nodes) {
    err = "empty mesh";
    return DGM_BAD_MESH;
  }
  m.nv = in.n_vertices;
  m.vx.resize(m.nv);
  m.vy.resize(m.nv);
  for (int64_t i = 0; i < m.nv; ++i) {
    m.vx[i] = in.vertices_xy[2 * i];
    m.vy[i] = in.vertices_xy[2 * i + 1];
  }
  int64_t ne = 0;
  for (int64_t b = 0; b < in.n_element_blocks; ++b) ne += in.element_block_count[b];
  m.ne = ne;
  if (in.element_xy) m.exy.assign(in.element_xy, in.element_xy + ne * 6);
  for (int64_t i = 0; i < 3 * ne; ++i)
    if (in.element_nodes[i] < 0 || in.element_nodes[i] >= m.nv) {
      err = "element node id out of range";
      return DGM_BAD_MESH;
    }

  // undirected node pair -> edge id
  std::unordered_map<uint64_t, int64_t, PairHash> id_of;
  id_of.reserve((size_t)(ne * 2 + 16));
  const uint64_t nvu = (uint64_t)m.nv;
  auto ukey = [nvu](int64_t a, int64_t b) {
    return a < b ? (uint64_t)a * nvu + (uint64_t)b : (uint64_t)b * nvu + (uint64_t)a;
  };

  // boundary blocks
  int64_t nb = 0;
  for (int64_t b = 0; b < in.n_boundary_blocks; ++b) nb += in.boundary_block_count[b];
  m.n_boundary_edges = nb;
  m.ed_vert.reserve((size_t)(nb + 2 * ne) * 2);
  m.ed_patch.reserve((size_t)(nb + 2 * ne));
  int64_t off = 0;
  for (int64_t b = 0; b < in.n_boundary_blocks; ++b) {
    const int32_t type = in.boundary_block_type[b];
    if (type != DGM_WALL && type != DGM_FARFIELD) {
      err = "boundary block type is neither wall nor farfield";
      return DGM_BAD_MESH;
    }
    m.patch_type.push_back(type);
    for (int64_t i = 0; i < in.boundary_block_count[b]; ++i, ++off) {
      const int64_t a = in.boundary_nodes[2 * off], c = in.boundary_nodes[2 * off + 1];
      if (a < 0 || a >= m.nv || c < 0 || c >= m.nv) {
        err = "boundary node id out of range";
        return DGM_BAD_MESH;
      }
      const int64_t id = (int64_t)m.ed_patch.size();
      m.ed_vert.push_back(a);
      m.ed_vert.push_back(c);
      m.ed_patch.push_back(b);
      id_of[ukey(a, c)] = id;  // a later duplicate line overwrites, as a map insert does
    }
  }

  // interior edges and element assignment, block by block
  m.el_vert.resize(3 * ne);
  m.el_edge.resize(3 * ne);
  m.el_nbr.assign(3 * ne, -1);
  std::vector<int64_t> slot0, slot1, lic0, lic1;  // grown with edges
  slot0.assign(nb, 0);
  slot1.assign(nb, -1);
  lic0.assign(nb, 0);
  lic1.assign(nb, -1);
  int64_t e0 = 0;
  for (int64_t blk = 0; blk < in.n_element_blocks; ++blk) {
    const int64_t cnt = in.element_block_count[blk];
    const int64_t* nodes = in.element_nodes + 3 * e0;
    for (int64_t t = 0; t < cnt; ++t)
      for (int i = 0; i < 3; ++i) {
        const int64_t a = nodes[3 * t + i], c = nodes[3 * t + (i + 1) % 3];
        auto ins = id_of.emplace(ukey(a, c), (int64_t)m.ed_patch.size());
        if (ins.second) {
          m.ed_vert.push_back(a);
          m.ed_vert.push_back(c);
          m.ed_patch.push_back(-1);
          slot0.push_back(0);
          slot1.push_back(0);
          lic0.push_back(0);
          lic1.push_back(0);
          m.internal_edges.push_back(ins.first->second);
        }
      }
    for (int64_t t = 0; t < cnt; ++t) {
      const int64_t el = e0 + t;
      bool boundary = false;
      for (int i = 0; i < 3; ++i) {
        const int64_t a = nodes[3 * t + i], c = nodes[3 * t + (i + 1) % 3];
        const int64_t id = id_of.at(ukey(a, c));
        m.el_vert[3 * el + i] = a;
        m.el_edge[3 * el + i] = id;
        if (slot1[id] == -1) {  // boundary edge
          slot0[id] = el;
          lic0[id] = i;
          boundary = true;
        } else if (a == m.ed_vert[2 * id]) {
          slot0[id] = el;
          lic0[id] = i;
        } else {
          slot1[id] = el;
          lic1[id] = i;
        }
      }
      (boundary ? m.boundary_elements : m.internal_elements).push_back(el);
    }
    e0 += cnt;
  }
  m.nedge = (int64_t)m.ed_patch.size();
  m.ed_elem.resize(2 * m.nedge);
  m.ed_lic.resize(2 * m.nedge);
  for (int64_t i = 0; i < m.nedge; ++i) {
    m.ed_elem[2 * i] = slot0[i];
    m.ed_elem[2 * i + 1] = slot1[i];
    m.ed_lic[2 * i] = lic0[i];
    m.ed_lic[2 * i + 1] = lic1[i];
  }
  // every boundary edge must belong to an element
  for (int64_t i = 0; i < nb; ++i) {
    bool owned = false;
    for (int k = 0; k < 3 && !owned; ++k) owned = m.el_edge[3 * slot0[i] + k] == i;
    if (!owned) {
      err = "boundary line not on any triangle";
      return DGM_BAD_MESH;
    }
  }

  auto vert_of_elem = [&m](int64_t e, int i, double& x, double& y) {
    if (!m.exy.empty()) {
      x = m.exy[(e * 3 + i) * 2];
      y = m.exy[(e * 3 + i) * 2 + 1];
    } else {
      x = m.vx[m.el_vert[3 * e + i]];
      y = m.vy[m.el_vert[3 * e + i]];
    }
  };
  // geometry
  m.ed_len.resize(m.nedge);
  m.ed_normal.resize(2 * m.nedge);
  m.ed_bc_normal.assign(2 * m.nedge, 0.0);
  for (int64_t i = 0; i < m.nedge; ++i) {
    double x1, y1, x2, y2;
    if (!m.exy.empty() && i >= nb) {  // interior edge of a periodic mesh: the left element's own copy
      const int64_t L = m.ed_elem[2 * i];
      const int lic = (int)m.ed_lic[2 * i];
      vert_of_elem(L, lic, x1, y1);
      vert_of_elem(L, (lic + 1) % 3, x2, y2);
    } else {
      x1 = m.vx[m.ed_vert[2 * i]];
      y1 = m.vy[m.ed_vert[2 * i]];
      x2 = m.vx[m.ed_vert[2 * i + 1]];
      y2 = m.vy[m.ed_vert[2 * i + 1]];
    }
    const double len = std::sqrt((x2 - x1) * (x2 - x1) + (y2 - y1) * (y2 - y1));
    m.ed_len[i] = len;
    m.ed_normal[2 * i] = (y2 - y1) / len;
    m.ed_normal[2 * i + 1] = -(x2 - x1) / len;
    if (i < nb) {  // src/spatial_disc/boundary_condition.rs:22-28
      const int64_t L = m.ed_elem[2 * i];
      const int lic = (int)m.ed_lic[2 * i];
      const double s = (m.el_vert[3 * L + lic] == m.ed_vert[2 * i]) ? 1.0 : -1.0;
      m.ed_bc_normal[2 * i] = s * m.ed_normal[2 * i];
      m.ed_bc_normal[2 * i + 1] = s * m.ed_normal[2 * i + 1];
    }
  }
  m.el_J.resize(ne);
  m.el_cr.resize(ne);
  m.el_hmin.resize(ne);
  m.el_metric.resize(4 * ne);
  for (int64_t e = 0; e < ne; ++e) {
    double x1, y1, x2, y2, x3, y3;
    vert_of_elem(e, 0, x1, y1);
    vert_of_elem(e, 1, x2, y2);
    vert_of_elem(e, 2, x3, y3);
    const double J = (x2 - x1) * (y3 - y1) - (x3 - x1) * (y2 - y1);
    if (!(J > 0.0)) {
      err = "triangle " + std::to_string(e) + " is not counter-clockwise";
      return DGM_BAD_MESH;
    }
    m.el_J[e] = J;
    m.el_metric[4 * e + 0] = (y3 - y1) / J;
    m.el_metric[4 * e + 1] = -(x3 - x1) / J;
    m.el_metric[4 * e + 2] = -(y2 - y1) / J;
    m.el_metric[4 * e + 3] = (x2 - x1) / J;
    const double a = std::sqrt((x2 - x1) * (x2 - x1) + (y2 - y1) * (y2 - y1));
    const double b = std::sqrt((x3 - x2) * (x3 - x2) + (y3 - y2) * (y3 - y2));
    const double c = std::sqrt((x1 - x3) * (x1 - x3) + (y1 - y3) * (y1 - y3));
    const double s = (a + b + c) / 2.0;
    m.el_cr[e] = a * b * c / (4.0 * std::sqrt(s * (s - a) * (s - b) * (s - c)));
    const double twice_area = J;  // 2 * (0.5 * J)
    m.el_hmin[e] = std::fmin(std::fmin(twice_area / a, twice_area / b), twice_area / c);
  }
  // neighbours
  for (int64_t e = 0; e < ne; ++e)
    for (int i = 0; i < 3; ++i) {
      const int64_t id = m.el_edge[3 * e + i];
      if (m.ed_elem[2 * id + 1] == -1) continue;
      m.el_nbr[3 * e + i] = (m.el_vert[3 * e + i] == m.ed_vert[2 * id]) ? m.ed_elem[2 * id + 1] : m.ed_elem[2 * id];
    }
  // level schedule: level(e) = 1 + max level of lower-numbered neighbours
  std::vector<int32_t> lvl(ne, 0);
  int32_t nl = 0;
  for (int64_t e = 0; e < ne; ++e) {
    int32_t l = 0;
    for (int i = 0; i < 3; ++i) {
      const int64_t n = m.el_nbr[3 * e + i];
      if (n >= 0 && n < e) l = std::max(l, lvl[n] + 1);
    }
    lvl[e] = l;
    nl = std::max(nl, l + 1);
  }
  m.level_offsets.assign(nl + 1, 0);
  for (int64_t e = 0; e < ne; ++e) m.level_offsets[lvl[e] + 1]++;
  for (int32_t l = 0; l < nl; ++l) m.level_offsets[l + 1] += m.level_offsets[l];
  m.level_elems.assign(ne, 0);
  std::vector<int32_t> fill(m.level_offsets.begin(), m.level_offsets.end() - 1);
  for (int64_t e = 0; e < ne; ++e) m.level_elems[fill[lvl[e]]++] = (int32_t)e;
  return DGM_OK;
}

}  // namespace dgm
