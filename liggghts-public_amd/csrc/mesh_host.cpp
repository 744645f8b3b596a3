// Host-side preparation of triangle meshes (see mesh_host.h).
//
// Element geometry: SurfaceMesh::calcEdgeVecLen / calcSurfaceNorm /
// calcEdgeNormals / calcObtuseAngleIndex (surface_mesh_I.h:361-447) and
// MultiNodeMesh::addElement (multi_node_mesh_I.h:139-172).
// Topology: SurfaceMesh::buildNeighbours (surface_mesh_I.h:454-560) with
// handleSharedEdge / handleCorner / checkNodeRecursive (:1035-1190),
// areCoplanar / coplanarNeighsOverlap / areCoplanarNodeNeighs (:844-1000),
// MultiNodeMesh::share2Nodes / nSharedNodes / nodesAreEqual (multi_node_mesh_I.h:246-355).
#include "mesh_host.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <fstream>
#include <sstream>

namespace dem {
namespace {

constexpr int kNeighMax = 5;  // SurfaceMesh<3,5>

bool comp_double(double a, double b, double prec) {  // math_extra_liggghts.h:536-547
  if (a == b) return true;
  if (b == 0) return a < prec && a > -prec;
  const double x = (a - b);
  return x < prec && x > -prec;
}

inline void sub3(const double *a, const double *b, double *r) {
  for (int k = 0; k < 3; k++) r[k] = a[k] - b[k];
}
inline double dot3(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
inline double mag3(const double *a) { return std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]); }
inline void cross3(const double *a, const double *b, double *r) {
  r[0] = a[1] * b[2] - a[2] * b[1];
  r[1] = a[2] * b[0] - a[0] * b[2];
  r[2] = a[0] * b[1] - a[1] * b[0];
}
inline void div3(double *v, double s) {
  const double sinv = 1. / s;
  for (int k = 0; k < 3; k++) v[k] = sinv * v[k];
}

bool nodes_equal(const HostMesh &m, const double *a, const double *b) {
  for (int k = 0; k < 3; k++)
    if (!comp_double(a[k], b[k], m.precision)) return false;
  return true;
}

void surf_props(HostMesh &m, int t) {
  const double *n = &m.node[9 * t];
  double *ev = &m.edgevec[9 * t], *el = &m.edgelen[3 * t], *sn = &m.snorm[3 * t];
  double *en = &m.edgenorm[9 * t];
  for (int i = 0; i < 3; i++) {
    sub3(&n[3 * ((i + 1) % 3)], &n[3 * i], &ev[3 * i]);
    el[i] = mag3(&ev[3 * i]);
    div3(&ev[3 * i], el[i]);
  }
  cross3(&ev[0], &ev[3], sn);
  if (mag3(sn) < 1e-15) {
    cross3(&ev[3], &ev[6], sn);
    if (mag3(sn) < 1e-15) {
      cross3(&ev[6], &ev[0], sn);
      if (mag3(sn) < 1e-15) {
        const double tmpvec[] = {1.1233, 2.123231, -3.3343434};
        cross3(&ev[0], tmpvec, sn);
        if (mag3(sn) < 1e-15) {
          cross3(&ev[3], tmpvec, sn);
          if (mag3(sn) < 1e-15) {
            const double tmpvec2[] = {1.1233, -2.123231, 3.3343434};
            cross3(&ev[0], tmpvec2, sn);
            if (mag3(sn) < 1e-15) cross3(&ev[3], tmpvec2, sn);
          }
        }
      }
    }
  }
  div3(sn, mag3(sn));
  for (int i = 0; i < 3; i++) {
    cross3(&ev[3 * i], sn, &en[3 * i]);
    if (mag3(&en[3 * i]) < 1e-15) {
      const int o = (i + 1) % 3;
      for (int k = 0; k < 3; k++) en[3 * i + k] = ev[3 * o + k];
    } else {
      div3(&en[3 * i], mag3(&en[3 * i]));
    }
  }
  // every node's test overwrites the element's index: the last node decides
  for (int i = 0; i < 3; i++) {
    const double d = dot3(&ev[3 * i], &ev[3 * ((i + 2) % 3)]);
    m.obtuse[t] = d > 0. ? i : -1;
  }
}

bool broad_phase(const HostMesh &m, int i, int j) {
  double d[3];
  sub3(&m.center[3 * i], &m.center[3 * j], d);
  const double radsum = m.rbound[i] + m.rbound[j];
  return !(d[0] * d[0] + d[1] * d[1] + d[2] * d[2] > radsum * radsum);
}

int n_shared_nodes(const HostMesh &m, int i, int j) {
  if (!broad_phase(m, i, j)) return 0;
  int ns = 0;
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++)
      if (nodes_equal(m, &m.node[9 * i + 3 * a], &m.node[9 * j + 3 * b])) ns++;
  return ns;
}

bool share_edge(const HostMesh &m, int i, int j, int &ie, int &je) {
  if (!broad_phase(m, i, j)) return false;
  int found = 0, in1 = -1, jn1 = -1, in2 = -1, jn2 = -1;
  for (int a = 0; a < 3 && found < 2; a++)
    for (int b = 0; b < 3; b++)
      if (nodes_equal(m, &m.node[9 * i + 3 * a], &m.node[9 * j + 3 * b])) {
        if (found == 0) {
          in1 = a;
          jn1 = b;
          found = 1;
        } else {
          in2 = a;
          jn2 = b;
          found = 2;
          break;
        }
      }
  if (found < 2) return false;
  ie = (in1 + in2 == 2) ? 2 : std::min(in1, in2);
  je = (jn1 + jn2 == 2) ? 2 : std::min(jn1, jn2);
  return true;
}

void shared_edge(HostMesh &m, int i, int ie, int j, int je, bool coplanar) {
  if (m.nneigh[i] < kNeighMax) m.neigh[kNeighMax * i + m.nneigh[i]] = j;
  if (m.nneigh[j] < kNeighMax) m.neigh[kNeighMax * j + m.nneigh[j]] = i;
  m.nneigh[i]++;
  m.nneigh[j]++;
  bool overlap = false;
  if (coplanar) {
    const double *pRef = &m.node[9 * i + 3 * ie];
    const double *edgeN = &m.edgenorm[9 * i + 3 * ie];
    double vI[3], vJ[3];
    sub3(&m.node[9 * i + 3 * ((ie + 2) % 3)], pRef, vI);
    sub3(&m.node[9 * j + 3 * ((je + 2) % 3)], pRef, vJ);
    overlap = dot3(vI, edgeN) * dot3(vJ, edgeN) > 0.;
  }
  if (!coplanar || overlap) {
    // the element with the larger index keeps the edge
    m.eact[3 * i + ie] = i < j ? 0 : 1;
    m.eact[3 * j + je] = i < j ? 1 : 0;
  } else {
    m.eact[3 * i + ie] = 0;
    m.eact[3 * j + je] = 0;
  }
}

int contains_node(const HostMesh &m, int t, const double *p) {
  for (int a = 0; a < 3; a++) {
    const double *q = &m.node[9 * t + 3 * a];
    if (comp_double(q[0], p[0], m.precision) && comp_double(q[1], p[1], m.precision) &&
        comp_double(q[2], p[2], m.precision))
      return a;
  }
  return -1;
}

struct CornerScan {
  std::vector<int> visited, has;
  std::vector<double> edges, ends;
  bool any_active = false;
};

void node_recursive(const HostMesh &m, int t, const double *p, CornerScan &c) {
  if (std::find(c.visited.begin(), c.visited.end(), t) != c.visited.end()) return;
  c.visited.push_back(t);
  const int a = contains_node(m, t, p);
  if (a < 0) return;
  c.has.push_back(t);
  const int am = (a + 2) % 3;
  for (int k = 0; k < 3; k++) c.edges.push_back(m.edgevec[9 * t + 3 * a + k]);
  for (int k = 0; k < 3; k++) c.edges.push_back(m.edgevec[9 * t + 3 * am + k]);
  for (int k = 0; k < 3; k++) c.ends.push_back(m.node[9 * t + 3 * ((a + 1) % 3) + k]);
  for (int k = 0; k < 3; k++) c.ends.push_back(m.node[9 * t + 3 * am + k]);
  if (m.eact[3 * t + a] || m.eact[3 * t + am]) c.any_active = true;
  for (int k = 0; k < std::min(m.nneigh[t], kNeighMax); k++) {
    const int nb = m.neigh[kNeighMax * t + k];
    if (nb < 0) return;
    node_recursive(m, nb, p, c);
  }
}

void corner(HostMesh &m, int t, int a, const double *lo, const double *hi) {
  const double *p = &m.node[9 * t + 3 * a];
  CornerScan c;
  node_recursive(m, t, p, c);
  // serial run: the subdomain is the box, padded by SMALL_DMBRDR (domain_I.h:65-86)
  for (int d = 0; d < 3; d++)
    if (!(p[d] >= lo[d] - 1.0e-8 && p[d] < hi[d] + 1.0e-8)) return;
  int maxId = -1;
  for (int id : c.has) maxId = std::max(maxId, id);
  bool colinear = false;
  const int ne = (int)c.edges.size() / 3;
  for (int i = 0; i < ne; i++)
    for (int j = i + 1; j < ne; j++)
      if (std::fabs(dot3(&c.edges[3 * i], &c.edges[3 * j])) > m.curvature &&
          !nodes_equal(m, &c.ends[3 * i], &c.ends[3 * j]))
        colinear = true;
  if (colinear || !c.any_active)
    m.cact[3 * t + a] = 0;
  else
    m.cact[3 * t + a] = (t == maxId) ? 1 : 0;
}

}  // namespace

bool build_host_mesh(HostMesh &m, int ntri, const double *nodes, const double *box_lo,
                     const double *box_hi, std::string &err) {
  m.node.clear();
  int nt = 0;
  for (int t = 0; t < ntri; t++) {
    const double *q = &nodes[9 * t];
    for (int k = 0; k < 9; k++)
      if (!std::isfinite(q[k])) {
        err = "non-finite mesh node";
        return false;
      }
    if (nodes_equal(m, &q[0], &q[3]) || nodes_equal(m, &q[3], &q[6]) || nodes_equal(m, &q[0], &q[6]))
      continue;  // degenerate element (MultiNodeMesh::addElement)
    m.node.insert(m.node.end(), q, q + 9);
    nt++;
  }
  if (nt == 0) {
    err = "mesh has no valid triangles";
    return false;
  }
  m.ntri = nt;
  m.center.assign(3 * nt, 0.0);
  m.rbound.assign(nt, 0.0);
  m.edgevec.assign(9 * nt, 0.0);
  m.edgelen.assign(3 * nt, 0.0);
  m.snorm.assign(3 * nt, 0.0);
  m.edgenorm.assign(9 * nt, 0.0);
  m.obtuse.assign(nt, -1);
  m.eact.assign(3 * nt, 1);
  m.cact.assign(3 * nt, 1);
  m.nneigh.assign(nt, 0);
  m.neigh.assign(kNeighMax * nt, -1);
  for (int t = 0; t < nt; t++) {
    const double *q = &m.node[9 * t];
    double avg[3] = {0., 0., 0.};
    for (int i = 0; i < 3; i++)
      for (int k = 0; k < 3; k++) avg[k] = q[3 * i + k] + avg[k];
    div3(avg, 3.0);
    double rb = 0.;
    for (int i = 0; i < 3; i++) {
      double v[3];
      sub3(avg, &q[3 * i], v);
      rb = std::max(rb, mag3(v));
    }
    for (int k = 0; k < 3; k++) m.center[3 * t + k] = avg[k];
    m.rbound[t] = rb;
    surf_props(m, t);
  }
  // each pair sharing an edge once, the later element as i
  for (int i = 0; i < nt; i++)
    for (int j = 0; j < i; j++) {
      int ie, je;
      if (share_edge(m, i, j, ie, je)) {
        const bool coplanar = std::fabs(dot3(&m.snorm[3 * i], &m.snorm[3 * j])) >= m.curvature;
        shared_edge(m, i, ie, j, je, coplanar);
      }
    }
  for (int t = 0; t < nt; t++)
    for (int a = 0; a < 3; a++) corner(m, t, a, box_lo, box_hi);
  // coplanar node neighbors: listed neighbor faces or elements sharing a node,
  // with |n_a . n_b| > curvature (SurfaceMesh::areCoplanarNodeNeighs)
  m.coplanar.assign(nt, {});
  for (int a = 0; a < nt; a++)
    for (int b = 0; b < nt; b++) {
      if (a == b) continue;
      bool are_neighs = false;
      for (int k = 0; k < std::min(m.nneigh[a], kNeighMax); k++)
        if (m.neigh[kNeighMax * a + k] == b) are_neighs = true;
      if (!are_neighs && n_shared_nodes(m, a, b) == 0) continue;
      if (std::fabs(dot3(&m.snorm[3 * a], &m.snorm[3 * b])) > m.curvature) m.coplanar[a].push_back(b);
    }
  return true;
}

void fill_tri_geom(const HostMesh &hm, int mesh_index, std::vector<TriGeom> &out) {
  for (int t = 0; t < hm.ntri; t++) {
    TriGeom g{};
    for (int k = 0; k < 9; k++) {
      g.n[k] = hm.node[9 * t + k];
      g.ev[k] = hm.edgevec[9 * t + k];
      g.vn[k] = 0.0;
    }
    for (int k = 0; k < 3; k++) {
      g.el[k] = hm.edgelen[3 * t + k];
      g.sn[k] = hm.snorm[3 * t + k];
      g.c[k] = hm.center[3 * t + k];
    }
    g.rb = hm.rbound[t];
    g.obtuse = hm.obtuse[t];
    int f = 0;
    for (int k = 0; k < 3; k++) {
      f |= (hm.eact[3 * t + k] ? 1 : 0) << k;
      f |= (hm.cact[3 * t + k] ? 1 : 0) << (3 + k);
    }
    g.flags = f;
    g.mesh = mesh_index;
    out.push_back(g);
  }
}

bool read_ascii_stl(const std::string &path, std::vector<double> &nodes, std::string &err) {
  std::ifstream f(path);
  if (!f) {
    err = "Cannot open mesh file " + path;
    return false;
  }
  nodes.clear();
  std::string line;
  while (std::getline(f, line)) {
    std::istringstream is(line);
    std::string w;
    if (!(is >> w)) continue;
    if (w == "vertex") {
      double x, y, z;
      if (!(is >> x >> y >> z)) {
        err = "bad vertex line in " + path;
        return false;
      }
      nodes.push_back(x);
      nodes.push_back(y);
      nodes.push_back(z);
    }
  }
  if (nodes.empty() || nodes.size() % 9 != 0) {
    err = "no triangles (or a partial one) in " + path + "; only ASCII STL is read";
    return false;
  }
  return true;
}

namespace {

constexpr double kGenPi = 3.14159265358979323846;  // M_PI

void add_tri(std::vector<double> &out, double ax, double ay, double az, double bx, double by, double bz, double cx,
             double cy, double cz) {
  const double v[9] = {ax, ay, az, bx, by, bz, cx, cy, cz};
  out.insert(out.end(), v, v + 9);
}

// points (r cos(i da), r sin(i da)), da = 2 pi / nseg (input_mesh_tri.cpp:820-828)
std::vector<double> ring(double r, int nseg) {
  std::vector<double> out(2 * (size_t)nseg);
  const double da = (2 * kGenPi) / (double)(nseg);
  for (int i = 0; i < nseg; ++i) {
    out[2 * i] = r * std::cos(i * da);
    out[2 * i + 1] = r * std::sin(i * da);
  }
  return out;
}

}  // namespace

bool generate_mesh(const std::string &type, const double *dv, int nseg, int mask, std::vector<double> &nodes,
                   std::string &err) {
  nodes.clear();
  if (type == "cube" || type == "box") {
    // face bits: top 1, bottom 2, front 4, back 8, left 16, right 32
    // (InputMeshTri::BoxMask, input_mesh_tri.h:66-73)
    const double x = dv[0] / 2.0, y = (type == "cube" ? dv[0] : dv[1]) / 2.0,
                 z = (type == "cube" ? dv[0] : dv[2]) / 2.0;
    if (mask & 16) {
      add_tri(nodes, -x, y, z, -x, y, -z, -x, -y, z);
      add_tri(nodes, -x, -y, z, -x, y, -z, -x, -y, -z);
    }
    if (mask & 32) {
      add_tri(nodes, x, y, -z, x, y, z, x, -y, -z);
      add_tri(nodes, x, -y, -z, x, y, z, x, -y, z);
    }
    if (mask & 1) {
      add_tri(nodes, x, y, z, -x, y, z, x, -y, z);
      add_tri(nodes, x, -y, z, -x, y, z, -x, -y, z);
    }
    if (mask & 2) {
      add_tri(nodes, -x, y, -z, x, y, -z, -x, -y, -z);
      add_tri(nodes, -x, -y, -z, x, y, -z, x, -y, -z);
    }
    if (mask & 4) {
      add_tri(nodes, x, -y, -z, x, -y, z, -x, -y, -z);
      add_tri(nodes, -x, -y, -z, x, -y, z, -x, -y, z);
    }
    if (mask & 8) {
      add_tri(nodes, x, y, z, x, y, -z, -x, y, z);
      add_tri(nodes, -x, y, z, x, y, -z, -x, y, -z);
    }
  } else if (type == "cylinder") {
    // face bits: top 1, bottom 2, side 4 (InputMeshTri::CylinderMask)
    const double z = dv[1] / 2.0;
    const std::vector<double> R = ring(dv[0], nseg);
    for (int i = 0; i < nseg; ++i) {
      const int p = i == 0 ? nseg - 1 : i - 1;
      const double xi = R[2 * i], yi = R[2 * i + 1], xp = R[2 * p], yp = R[2 * p + 1];
      if (mask & 1) add_tri(nodes, 0, 0, z, xi, yi, z, xp, yp, z);
      if (mask & 2) add_tri(nodes, 0, 0, -z, xi, yi, -z, xp, yp, -z);
      if (mask & 4) {
        add_tri(nodes, xi, yi, z, xp, yp, z, xi, yi, -z);
        add_tri(nodes, xp, yp, z, xp, yp, -z, xi, yi, -z);
      }
    }
  } else if (type == "pipe") {
    // face bits: inner top 1, outer top 2, inner bottom 4, outer bottom 8,
    // inner side 16, outer side 32 (InputMeshTri::PipeMask)
    if (!(dv[1] > dv[0])) {
      err = "fix mesh/surface generate pipe: outer_radius must exceed inner_radius";
      return false;
    }
    const double z = dv[2] / 2.0;
    const std::vector<double> I = ring(dv[0], nseg), O = ring(dv[1], nseg);
    for (int i = 0; i < nseg; ++i) {
      const int p = i == 0 ? nseg - 1 : i - 1;
      const double ix = I[2 * i], iy = I[2 * i + 1], ipx = I[2 * p], ipy = I[2 * p + 1];
      const double ox = O[2 * i], oy = O[2 * i + 1], opx = O[2 * p], opy = O[2 * p + 1];
      if (mask & 1) add_tri(nodes, 0, 0, z, ix, iy, z, ipx, ipy, z);
      if (mask & 2) {
        add_tri(nodes, ix, iy, z, ipx, ipy, z, ox, oy, z);
        add_tri(nodes, ipx, ipy, z, opx, opy, z, ox, oy, z);
      }
      if (mask & 4) add_tri(nodes, 0, 0, -z, ix, iy, -z, ipx, ipy, -z);
      if (mask & 8) {
        add_tri(nodes, ix, iy, -z, ipx, ipy, -z, ox, oy, -z);
        add_tri(nodes, ipx, ipy, -z, opx, opy, -z, ox, oy, -z);
      }
      if (mask & 16) {
        add_tri(nodes, ix, iy, z, ipx, ipy, z, ix, iy, -z);
        add_tri(nodes, ipx, ipy, z, ipx, ipy, -z, ix, iy, -z);
      }
      if (mask & 32) {
        add_tri(nodes, ox, oy, z, opx, opy, z, ox, oy, -z);
        add_tri(nodes, opx, opy, z, opx, opy, -z, ox, oy, -z);
      }
    }
  } else if (type == "disk") {
    const std::vector<double> R = ring(dv[0], nseg);
    for (int i = 0; i < nseg; ++i) {
      const int p = i == 0 ? nseg - 1 : i - 1;
      add_tri(nodes, 0, 0, 0, R[2 * i], R[2 * i + 1], 0, R[2 * p], R[2 * p + 1], 0);
    }
  } else if (type == "plane") {
    const double a = dv[0] / 2.0;
    add_tri(nodes, a, a, 0, -a, a, 0, a, -a, 0);
    add_tri(nodes, a, -a, 0, -a, a, 0, -a, -a, 0);
  } else {
    err = "unknown generator type '" + type + "'";
    return false;
  }
  if (nodes.empty()) {
    err = "fix mesh/surface generate " + type + ": the mask selects no face";
    return false;
  }
  return true;
}

}  // namespace dem
