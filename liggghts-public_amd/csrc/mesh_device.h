// Device functions for triangle-mesh walls, FP64 throughout (the overlap of a
// sphere with a triangle is a difference of nearly equal lengths). Included by
// dem_kernels.hip, compiled with -ffp-contract=off so every rounding matches
// the CPU reference arithmetic.
#pragma once

#include "dem_device.h"

namespace dem {

constexpr double kSmallTrimesh = 1.e-10;  // tri_mesh_I.h:47-50
constexpr double kLargeTrimesh = 1000000.;

__device__ __forceinline__ void d3_sub(const double *a, const double *b, double *r) {
  r[0] = a[0] - b[0];
  r[1] = a[1] - b[1];
  r[2] = a[2] - b[2];
}
__device__ __forceinline__ double d3_dot(const double *a, const double *b) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}
__device__ __forceinline__ double d3_mag(const double *a) {
  return sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
}
__device__ __forceinline__ void d3_cross(const double *a, const double *b, double *r) {
  r[0] = a[1] * b[2] - a[2] * b[1];
  r[1] = a[2] * b[0] - a[0] * b[2];
  r[2] = a[0] * b[1] - a[1] * b[0];
}
__device__ __forceinline__ void d3_div(double *v, double sc) {
  const double sinv = 1. / sc;
  v[0] = sinv * v[0];
  v[1] = sinv * v[1];
  v[2] = sinv * v[2];
}

// three scalars, so the vector stays in registers when passed by reference
struct Vec3 {
  double x, y, z;
};

// |p - cp| and delta = cp - p (TriMesh::calcDist)
__device__ __forceinline__ double tri_calc_dist(const double *p, const double *cp, Vec3 &delta) {
  delta.x = cp[0] - p[0];
  delta.y = cp[1] - p[1];
  delta.z = cp[2] - p[2];
  return sqrt((p[0] - cp[0]) * (p[0] - cp[0]) + (p[1] - cp[1]) * (p[1] - cp[1]) +
              (p[2] - cp[2]) * (p[2] - cp[2]));
}

// barycentric coordinates held as three scalars (an array passed by pointer
// through the switch below ends up in scratch memory)
struct Bary {
  double b0, b1, b2;
  template <int K> __device__ __forceinline__ void set(double v) {
    if (K == 0) b0 = v;
    else if (K == 1) b1 = v;
    else b2 = v;
  }
  template <int K> __device__ __forceinline__ double get() const {
    return K == 0 ? b0 : (K == 1 ? b1 : b2);
  }
};

__device__ __forceinline__ bool edge_on(const TriGeom &t, int k) { return (t.flags >> k) & 1; }
__device__ __forceinline__ bool corner_on(const TriGeom &t, int k) {
  return (t.flags >> (3 + k)) & 1;
}

// TriMesh::resolveEdgeContactBary (tri_mesh_I.h:120-165); the edge index is a
// template argument so the barycentric array is indexed by constants
template <int iEdge>
__device__ __forceinline__ double tri_edge(const TriGeom &t, const double *p, Vec3 &delta, Bary &bary) {
  constexpr int ip = (iEdge + 1) % 3, ipp = (iEdge + 2) % 3;
  const double *n0 = &t.n[3 * iEdge];
  double nodeToP[3];
  d3_sub(p, n0, nodeToP);
  const double distFromNode = d3_dot(nodeToP, &t.ev[3 * iEdge]);
  const double len = t.el[iEdge];
  if (distFromNode < -kSmallTrimesh) {
    if (!corner_on(t, iEdge)) return kLargeTrimesh;
    const double d = tri_calc_dist(p, n0, delta);
    bary.set<iEdge>(1.);
    bary.set<ip>(0.);
    bary.set<ipp>(0.);
    return d;
  } else if (distFromNode > len + kSmallTrimesh) {
    if (!corner_on(t, ip)) return kLargeTrimesh;
    const double d = tri_calc_dist(p, &t.n[3 * ip], delta);
    bary.set<iEdge>(0.);
    bary.set<ip>(1.);
    bary.set<ipp>(0.);
    return d;
  }
  if (!edge_on(t, iEdge)) return kLargeTrimesh;
  double cp[3];
  for (int k = 0; k < 3; k++) cp[k] = n0[k] + distFromNode * t.ev[3 * iEdge + k];
  const double d = tri_calc_dist(p, cp, delta);
  bary.set<ipp>(0.);
  bary.set<iEdge>(1. - distFromNode / len);
  bary.set<ip>(1. - bary.get<iEdge>());
  return d;
}

// TriMesh::resolveCornerContactBary (tri_mesh_I.h:169-240)
template <int iNode>
__device__ __forceinline__ double tri_corner(const TriGeom &t, bool obtuse, const double *p,
                                    Vec3 &delta, Bary &bary) {
  constexpr int ip = (iNode + 1) % 3, ipp = (iNode + 2) % 3;
  const double *n = &t.n[3 * iNode];
  if (obtuse) {
    double nodeToP[3], cp[3];
    d3_sub(p, n, nodeToP);
    double distFromNode = d3_dot(nodeToP, &t.ev[3 * ipp]);
    if (distFromNode < kSmallTrimesh) {
      if (distFromNode > -t.el[ipp]) {
        if (!edge_on(t, ipp)) return kLargeTrimesh;
        for (int k = 0; k < 3; k++) cp[k] = n[k] + distFromNode * t.ev[3 * ipp + k];
        bary.set<ip>(0.);
        bary.set<iNode>(1. + distFromNode / t.el[ipp]);
        bary.set<ipp>(1. - bary.get<iNode>());
        return tri_calc_dist(p, cp, delta);
      }
      if (!corner_on(t, ipp)) return kLargeTrimesh;
      bary.set<ipp>(1.);
      bary.set<iNode>(0.);
      bary.set<ip>(0.);
      return tri_calc_dist(p, &t.n[3 * ipp], delta);
    }
    distFromNode = d3_dot(nodeToP, &t.ev[3 * iNode]);
    if (distFromNode > -kSmallTrimesh) {
      if (distFromNode < t.el[iNode]) {
        if (!edge_on(t, iNode)) return kLargeTrimesh;
        for (int k = 0; k < 3; k++) cp[k] = n[k] + distFromNode * t.ev[3 * iNode + k];
        bary.set<ipp>(0.);
        bary.set<iNode>(1. - distFromNode / t.el[iNode]);
        bary.set<ip>(1. - bary.get<iNode>());
        return tri_calc_dist(p, cp, delta);
      }
      if (!corner_on(t, ip)) return kLargeTrimesh;
      bary.set<ip>(1.);
      bary.set<iNode>(0.);
      bary.set<ipp>(0.);
      return tri_calc_dist(p, &t.n[3 * ip], delta);
    }
  }
  if (!corner_on(t, iNode)) return kLargeTrimesh;
  bary.set<iNode>(1.);
  bary.set<ip>(0.);
      bary.set<ipp>(0.);
  return tri_calc_dist(p, n, delta);
}

// TriMesh::resolveTriSphereContactBary (tri_mesh_I.h:65-116), barycentric
// coordinates by MathExtraLiggghts::calcBaryTriCoords (math_extra_liggghts.h:559-570)
__device__ __forceinline__ double tri_sphere(const TriGeom &t, double precision, double r, const double *p,
                                    Vec3 &delta, Bary &bary, int &barysign) {
  double n0p[3];
  d3_sub(p, &t.n[0], n0p);
  const double a = d3_dot(n0p, &t.ev[0]);
  const double b = d3_dot(n0p, &t.ev[6]);
  const double c = d3_dot(&t.ev[0], &t.ev[6]);
  const double oneMinCSqr = 1 - c * c;
  bary.b1 = (a - b * c) / (t.el[0] * oneMinCSqr);
  bary.b2 = (a * c - b) / (t.el[2] * oneMinCSqr);
  bary.b0 = 1. - bary.b1 - bary.b2;
  const double invlen = 1. / (2. * t.rb);
  const int bs = (bary.b0 > -precision * invlen) + 2 * (bary.b1 > -precision * invlen) +
                 4 * (bary.b2 > -precision * invlen);
  barysign = bs;
  double d;
  switch (bs) {
    case 1: d = tri_corner<0>(t, t.obtuse == 0, p, delta, bary); break;
    case 2: d = tri_corner<1>(t, t.obtuse == 1, p, delta, bary); break;
    case 3: d = tri_edge<0>(t, p, delta, bary); break;
    case 4: d = tri_corner<2>(t, t.obtuse == 2, p, delta, bary); break;
    case 5: d = tri_edge<2>(t, p, delta, bary); break;
    case 6: d = tri_edge<1>(t, p, delta, bary); break;
    case 7: {
      const double dNorm = d3_dot(t.sn, n0p);
      double cs[3];
      for (int k = 0; k < 3; k++) cs[k] = p[k] - t.sn[k] * dNorm;
      d = tri_calc_dist(p, cs, delta);
      break;
    }
    default: d = 1.; break;
  }
  return d - r;
}

// squared distance from p to the triangle (closest point after Ericson,
// Real-Time Collision Detection 5.1.5); selects candidate triangles only
__device__ inline double tri_point_dist2(const double *a, const double *b, const double *c,
                                         const double *p) {
  double ab[3], ac[3], ap[3], q[3];
  d3_sub(b, a, ab);
  d3_sub(c, a, ac);
  d3_sub(p, a, ap);
  const double d1 = d3_dot(ab, ap), d2 = d3_dot(ac, ap);
  if (d1 <= 0. && d2 <= 0.) {
    q[0] = a[0]; q[1] = a[1]; q[2] = a[2];
  } else {
    double bp[3];
    d3_sub(p, b, bp);
    const double d3 = d3_dot(ab, bp), d4 = d3_dot(ac, bp);
    if (d3 >= 0. && d4 <= d3) {
      q[0] = b[0]; q[1] = b[1]; q[2] = b[2];
    } else {
      const double vc = d1 * d4 - d3 * d2;
      if (vc <= 0. && d1 >= 0. && d3 <= 0.) {
        const double v = d1 / (d1 - d3);
        for (int k = 0; k < 3; k++) q[k] = a[k] + v * ab[k];
      } else {
        double cp[3];
        d3_sub(p, c, cp);
        const double d5 = d3_dot(ab, cp), d6 = d3_dot(ac, cp);
        if (d6 >= 0. && d5 <= d6) {
          q[0] = c[0]; q[1] = c[1]; q[2] = c[2];
        } else {
          const double vb = d5 * d2 - d1 * d6;
          if (vb <= 0. && d2 >= 0. && d6 <= 0.) {
            const double w = d2 / (d2 - d6);
            for (int k = 0; k < 3; k++) q[k] = a[k] + w * ac[k];
          } else {
            const double va = d3 * d6 - d5 * d4;
            if (va <= 0. && (d4 - d3) >= 0. && (d5 - d6) >= 0.) {
              const double w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
              for (int k = 0; k < 3; k++) q[k] = b[k] + w * (c[k] - b[k]);
            } else {
              const double denom = 1. / (va + vb + vc);
              const double v = vb * denom, w = vc * denom;
              for (int k = 0; k < 3; k++) q[k] = a[k] + ab[k] * v + ac[k] * w;
            }
          }
        }
      }
    }
  }
  double d[3];
  d3_sub(p, q, d);
  return d3_dot(d, d);
}

__device__ __forceinline__ void quatquat(const double *a, const double *b, double *c) {
  c[0] = a[0] * b[0] - a[1] * b[1] - a[2] * b[2] - a[3] * b[3];
  c[1] = a[0] * b[1] + b[0] * a[1] + a[2] * b[3] - a[3] * b[2];
  c[2] = a[0] * b[2] + b[0] * a[2] + a[3] * b[1] - a[1] * b[3];
  c[3] = a[0] * b[3] + b[0] * a[3] + a[1] * b[2] - a[2] * b[1];
}

// v <- q v q* (MathExtraLiggghts::vec_quat_rotate, math_extra_liggghts.h:453-470)
__device__ __forceinline__ void vec_quat_rotate(double *v, const double *q) {
  const double vq[4] = {0., v[0], v[1], v[2]};
  const double qc[4] = {q[0], -q[1], -q[2], -q[3]};
  double t[4], r[4];
  quatquat(q, vq, t);
  quatquat(t, qc, r);
  v[0] = r[1];
  v[1] = r[2];
  v[2] = r[3];
}

// Edge vectors and lengths, surface normal and obtuse index from the nodes
// (SurfaceMesh::calcEdgeVecLen / calcSurfaceNorm / calcObtuseAngleIndex,
// surface_mesh_I.h:361-447); the obtuse index keeps the last node's test.
__device__ inline void tri_surf_props(TriGeom &t) {
  for (int i = 0; i < 3; i++) {
    d3_sub(&t.n[3 * ((i + 1) % 3)], &t.n[3 * i], &t.ev[3 * i]);
    t.el[i] = d3_mag(&t.ev[3 * i]);
    d3_div(&t.ev[3 * i], t.el[i]);
  }
  double *sn = t.sn;
  d3_cross(&t.ev[0], &t.ev[3], sn);
  if (d3_mag(sn) < 1e-15) {
    d3_cross(&t.ev[3], &t.ev[6], sn);
    if (d3_mag(sn) < 1e-15) {
      d3_cross(&t.ev[6], &t.ev[0], sn);
      if (d3_mag(sn) < 1e-15) {
        const double tmpvec[] = {1.1233, 2.123231, -3.3343434};
        d3_cross(&t.ev[0], tmpvec, sn);
        if (d3_mag(sn) < 1e-15) {
          d3_cross(&t.ev[3], tmpvec, sn);
          if (d3_mag(sn) < 1e-15) {
            const double tmpvec2[] = {1.1233, -2.123231, 3.3343434};
            d3_cross(&t.ev[0], tmpvec2, sn);
            if (d3_mag(sn) < 1e-15) d3_cross(&t.ev[3], tmpvec2, sn);
          }
        }
      }
    }
  }
  d3_div(sn, d3_mag(sn));
  int ob = -1;
  for (int i = 0; i < 3; i++) {
    const double dot = d3_dot(&t.ev[3 * i], &t.ev[3 * ((i + 2) % 3)]);
    ob = dot > 0. ? i : -1;
  }
  t.obtuse = ob;
}

}  // namespace dem
