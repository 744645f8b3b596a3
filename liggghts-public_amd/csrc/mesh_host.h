// Host-side preparation of fix mesh/surface triangle meshes: element
// geometry and topology (edge and corner activity, neighbor faces, coplanar
// node neighbors), computed once when a mesh is added.
#pragma once

#include <string>
#include <vector>

#include "dem_device.h"

namespace dem {

struct HostMesh {
  int ntri = 0;
  std::vector<double> node;      // 9 per triangle
  std::vector<double> center;    // 3
  std::vector<double> rbound;    // 1
  std::vector<double> edgevec;   // 9
  std::vector<double> edgelen;   // 3
  std::vector<double> snorm;     // 3
  std::vector<double> edgenorm;  // 9
  std::vector<int> obtuse;       // 1
  std::vector<unsigned char> eact, cact;  // 3 each
  std::vector<int> nneigh, neigh;         // neighbor faces (5 slots)
  std::vector<std::vector<int>> coplanar;  // coplanar node neighbors per triangle
  double precision = 1e-8;                 // EPSILON_PRECISION (multi_node_mesh.h:59)
  double curvature = 1. - 0.00001;         // 1 - EPSILON_CURVATURE (surface_mesh.h:61)
  int wall_type = 1;
  int model = 0;  // contact-model bits
  bool moving = false;
  int mover = 0;  // dem::kMover* of its fix move/mesh (0: none)
  double origin[3] = {0, 0, 0}, axis[3] = {0, 0, 1}, omega = 0.0;  // rotate
  double vel[3] = {0, 0, 0};                                       // linear: velocity
  double ampl[3] = {0, 0, 0}, womega = 0.0;                        // wiggle
  // viblin (unit axis in `axis`) and vibrot (origin, unit axis): the order
  // and per-term amplitude, phase, angular frequency 2 pi / period
  int vord = 0;
  double vampl[30] = {}, vphi[30] = {}, vomega[30] = {};
  bool stress = false;          // fix mesh/surface/stress: total force and torque tracked
  double p_ref[3] = {0, 0, 0};  // its reference_point
};

// Builds a mesh from ntri*9 node coordinates: drops elements with coinciding
// nodes (MultiNodeMesh::addElement), computes per-element geometry and the
// topology. box_lo/box_hi bound the simulation box (corner activity is only
// decided for nodes inside it, as Domain::is_in_subdomain does).
bool build_host_mesh(HostMesh &m, int ntri, const double *nodes, const double *box_lo,
                     const double *box_hi, std::string &err);

// Fills one device record per triangle (mesh index m, global index offset
// is the caller's business).
void fill_tri_geom(const HostMesh &hm, int mesh_index, std::vector<TriGeom> &out);

// Reads the triangles of an ASCII STL file (input_mesh_tri.cpp:318).
bool read_ascii_stl(const std::string &path, std::vector<double> &nodes, std::string &err);

// fix mesh/surface ... generate <type> <values>: the triangles of the mesh
// generators (InputMeshTri::generate_box / cylinder / pipe / disk / plane,
// input_mesh_tri.cpp:688-1107), centred on the origin, in the generator's
// triangle and node order (the order fixes the element ids the topology rules
// use). dv: cube size | box xsize ysize zsize | cylinder radius zsize | pipe
// inner_radius outer_radius zsize | disk radius | plane size; nseg: number of
// segments (cylinder, pipe, disk); mask: the face bits (cube, box, cylinder,
// pipe).
bool generate_mesh(const std::string &type, const double *dv, int nseg, int mask, std::vector<double> &nodes,
                   std::string &err);

}  // namespace dem
