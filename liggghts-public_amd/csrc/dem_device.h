// Device-side data layout and kernel launchers of the DEM step library.
// Host code (dem_system.cpp) includes this header; the kernels live in
// dem_kernels.hip. No torch or Python types cross this boundary.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dem {

constexpr int kMaxWalls = 6;
constexpr int kMaxTypes = 8;
constexpr int kBlock = 256;
// A neighbor entry is the neighbor's index plus two bits: kFlagBit is the
// reference's contact flag being nonzero (the pair's history is valid and is
// carried over at the next build), kCarryBit that the build carried the pair
// over (the build stores the flag as 1, the normal-model bit, which a
// hertz/stiffness pair keeps in the surfacesClose band and others set anyway).
constexpr uint32_t kFlagBit = 0x80000000u;
constexpr uint32_t kCarryBit = 0x40000000u;
constexpr uint32_t kIndexMask = 0x3fffffffu;
constexpr int kMetaCountMask = 0xffff;       // meta[i]: low 16 bits = list length
constexpr int kMetaWallShift = 16;           // bits 16..21: wall history present
constexpr int kMaxMeshes = 4;
constexpr int kMeshSlots = 8;       // contact-history slots per particle and mesh
constexpr int kMaxMeshCand = 32;    // candidate triangles per particle (all meshes)
// omega.w of a particle holds type + (instance << kTypeBits): independent
// simulations batched in one system each own a cell grid and a material block
constexpr int kTypeBits = 8;
constexpr int kTypeMask = (1 << kTypeBits) - 1;
constexpr int kMaxInstances = 1 << 16;
// Tile shapes the step kernel is compiled for (threads per block, neighbor
// entries whose loads are issued together, and whether the touching entries
// are compacted over the wavefront before the contact law: 1 group by group,
// 2 over whole list rows; 3: per lane, with the shear histories loaded in the
// broad phase); variant 0 is the default. The sibling contact models run in
// variants 0, 5 and 7.
constexpr int kNumStepVariants = 11;
constexpr int kStepVariantBlock[kNumStepVariants] = {256, 256, 256, 128, 64, 256, 256, 256, 256, 256, 128};
constexpr int kStepVariantGroup[kNumStepVariants] = {4, 2, 8, 4, 4, 4, 8, 4, 8, 4, 4};
constexpr int kStepVariantWave[kNumStepVariants] = {0, 0, 0, 0, 0, 1, 1, 2, 2, 3, 3};

// One triangle of a fix mesh/surface, FP64 (TriMesh element properties).
struct TriGeom {
  double n[9];   // nodes
  double ev[9];  // unit edge vectors, edge k from node k to node k+1
  double el[3];  // edge lengths
  double sn[3];  // unit surface normal
  double c[3];   // centre
  double rb;     // bounding radius
  double vn[9];  // node velocities
  int obtuse;    // node index with the obtuse angle, or -1
  int flags;     // bits 0-2 edge active, bits 3-5 corner active
  int mesh;      // mesh index
  int pad;
};

// Scalar run parameters, passed by value to every kernel.
struct DevParams {
  double dt, dtf, dtfrotate;
  double triggersq;        // 0.25 * skin * skin (neighbor.cpp:290)
  int neigh_every;         // neigh_modify every (>= 1)
  int neigh_check;         // neigh_modify check yes (1) / no (0)
  double skin, cdf;
  double cdm;              // contactDistanceMultiplier = cdf * cdf (pair_gran_base.h:224)
  int close_band;          // cdf > 1: pairs within cdf * radsum that do not touch
                           // take the surfacesClose branch
  double lo[3], L[3], halfL[3];
  int per[3];
  int lost_faces;          // f faces of non-periodic dimensions (bit 2d lower, 2d+1 upper):
                           // an atom beyond one at a build is lost
  int pair_model;          // contact-model bits of pair_style gran (kModel*)
  int generic;             // some pair, wall or mesh model is not Hertz / history
  int hist_w;              // history values per pair / wall contact: 3 shear, 6 with
                           // the epsd2 rolling torque (generic kernels only)
  double char_vel;         // characteristicVelocity of model hooke
  int has_grav, grav_first;
  double gacc[3];
  int nwalls;
  int wall_axis[kMaxWalls];
  double wall_pos[kMaxWalls];
  int wall_type[kMaxWalls];
  int wall_model[kMaxWalls];  // contact-model bits of each fix wall/gran
  double cutneighmax;
  int ntp1;                // ntypes + 1 (row length of the material tables)
  // triangle meshes
  int nmesh;
  int mesh_tri0[kMaxMeshes], mesh_ntri[kMaxMeshes];
  int mesh_wtype[kMaxMeshes], mesh_model[kMaxMeshes], mesh_moving[kMaxMeshes];
  int mesh_stress;                 // bit m: mesh m is a mesh/surface/stress
  double mesh_pref[kMaxMeshes][3];  // its reference_point
  double mesh_precision;   // EPSILON_PRECISION of the mesh (multi_node_mesh.h:59)
  // cell grid of one instance; instance k owns global cells
  // [k * ncell_local, (k + 1) * ncell_local)
  int nc[3];
  double cinv[3];
  int ninst, ncell_local;
  int mat_stride;          // ntp1 * ntp1: material block of one instance
};

// Contact-model bits (one int per pair style / wall fix): bit 0 rolling cdt,
// 1 tangential damping, 2 limitForce, 3 torsionTorque, 4-5 the normal model,
// 6 tangential no_history, 7 cohesion sjkr, 8 ktToKnUser (hooke),
// 9 absolute_damping (hooke/stiffness), 10 tangential_reduce (sjkr),
// 11 rolling_friction epsd2 (elastic rolling spring with a torque history).
constexpr int kModelNormalShift = 4;
constexpr int kNormalHertz = 0, kNormalHooke = 1, kNormalHertzStiffness = 2,
              kNormalHookeStiffness = 3;
constexpr int kModelNoHistory = 1 << 6, kModelSjkr = 1 << 7, kModelKtToKn = 1 << 8,
              kModelAbsDamping = 1 << 9, kModelTangReduce = 1 << 10, kModelEpsd2 = 1 << 11;
// bits that need the generic contact law (anything beyond Hertz / history)
constexpr int kModelGenericMask = ~0xf;

// Mixed material tables, [ntp1*ntp1] each, FP64 and FP32 copies. kn, kt,
// gamman, gammat, cohesion energy density and log(e) feed the hooke,
// stiffness and sjkr models.
constexpr int kNumMatTables = 11;
struct MatTables {
  double *Yeff, *Geff, *beta, *mu, *mur, *kn, *kt, *gn, *gt, *ced, *loge;
  float *fYeff, *fGeff, *fbeta, *fmu, *fmur, *fkn, *fkt, *fgn, *fgt, *fced, *floge;
};

// Control block shared between the step kernels and the host driver.
struct Ctrl {
  long long stop_at;     // first step whose kernel must not run (rebuild pending)
  int overflow;          // widest list seen when it exceeded the capacity
  int maxcount;          // widest list at the last build
  unsigned long long ncontact_full;  // touching full-list entries at the last counted pass
  unsigned long long nwall;          // touching wall contacts at the last counted pass
  unsigned long long npairs_full;    // full-list entries at the last build
  int bad;               // non-finite state detected
  int mesh_overflow;     // candidate triangles of a particle beyond kMaxMeshCand
  int slot_overflow;     // simultaneous mesh contacts of a particle beyond kMeshSlots
  int pad2;
  unsigned long long nmesh_cand;  // (particle, triangle) candidates at the last build
  long long step_base, ago_base;  // step index and neighbor age of a graph replay's first launch
  // mesh/surface/stress: total force and torque (about the mesh's reference
  // point) the particles exert on each mesh, at the last counted pass
  double mesh_force[kMaxMeshes][6];
  unsigned long long nlost;  // atoms beyond an f face at the last build
};

// Particle state read by neighbors: ping-pong buffered across steps.
struct StateBuf {
  double4 *pos;   // x, y, z, radius
  double4 *vel;   // vx, vy, vz, rmass
  void *omg;      // double4 (FP64 mode) or float4 (mixed): wx, wy, wz, type
};

// Per-particle data touched only by the particle's own thread: updated in
// place by the step kernel, swapped only by a rebuild.
struct AuxBuf {
  int *tag;
  int *orig;      // caller index
  int *meta;      // list length | wall-history bits
  void *whist;    // wall history, [wall][hist_w][cap], double or float
  int *mslot;     // mesh history slots: triangle index or -1, [mesh][slot][cap]
  void *mhist;    // mesh history, [mesh][slot][hist_w][cap] (shear, then the epsd2 rolling torque)
};

// Mesh data used by the step kernel.
struct MeshStep {
  TriGeom *tri;
  const int *copl_off, *copl_idx;  // coplanar node-neighbor lists per triangle
  const int *cand;                 // [kMaxMeshCand][cap] candidate triangles, ascending
  const int *ncand;                // [cap]
};

struct ListBuf {
  uint32_t *nbr;  // [W][cap] neighbor index | contact flag
  void *hist;     // [W][cap][hist_rec] tangential (+ rolling) history, double or float
};

// The history values of one list entry sit together in a record of
// hist_rec(hist_w) values (4 for 3, 8 for 6: the shear in values 0-2, the
// epsd2 rolling torque in 4-6), the records of one list row consecutive, so
// one entry's history is one aligned vector load and store.
__host__ __device__ constexpr int hist_rec(int hist_w) { return hist_w <= 4 ? 4 : 8; }
// offset of history value d (0-2 shear, 3-5 rolling torque) in its record
__host__ __device__ constexpr int hist_off(int d) { return d < 3 ? d : d + 1; }
// first value of the record of list entry (slot k, particle i)
__host__ __device__ inline size_t hist_at(size_t k, size_t i, size_t cap, size_t rec) {
  return (k * cap + i) * rec;
}

// Rigid rotation of the moving meshes for one step.
// fix move/mesh movement types
constexpr int kMoverRotate = 1, kMoverLinear = 2, kMoverWiggle = 3, kMoverVibLin = 4,
              kMoverVibRot = 5;
// motion table of the time-dependent movers: per run step and mesh, the
// incremental rotation quaternion (4), the angular velocity vector (3) and the
// velocity factor of a translation (1), evaluated on the host
constexpr int kMotionStride = 8;

struct MeshMoveArgs {
  TriGeom *tri;               // geometry of the step before
  TriGeom *dst;               // where the moved geometry goes (nullptr: in place)
  const double *node_last;
  int ntri;
  int moving_mask;            // bit m: mesh m moves
  int kind[kMaxMeshes];       // kMover* of mesh m
  double dQ[kMaxMeshes][4];   // rotate: incremental rotation quaternion
  double origin[kMaxMeshes][3];
  double omegaVec[kMaxMeshes][3];
  int trans[kMaxMeshes];
  double tvec[kMaxMeshes][3];  // linear: velocity; wiggle: amplitude; viblin: unit axis
  const double *motion;        // [run step][kMaxMeshes][kMotionStride] (wiggle, viblin, vibrot)
  double dt;
  double triggersq;
  long long step;             // the step the moved geometry belongs to
  int guard;                  // return if step >= stop_at
  int decide;                 // raise stop_at = step+1 when a node moved > skin/2
  Ctrl *ctrl;
};

struct StepArgs {
  DevParams prm;
  MatTables mat;
  int n, cap, W;
  int mixed;
  StateBuf in, out;
  AuxBuf aux;
  ListBuf list;
  MeshStep mesh;
  double *xhold;     // [3][cap]
  double *fout;      // [6][cap] f and torque when store_f
  Ctrl *ctrl;
  unsigned long long *inst_stats;  // [2][ninst] touching pairs (full list) and wall contacts, or null
  long long step;
  long long ago_next;  // steps since the last build at the step this launch integrates to
  int shearupdate, do_final, do_initial, store_f, count;
  // tile shape (kStepVariant*) and the number of blocks that walk the
  // particle tiles (all tiles: one each; fewer: a persistent grid); with
  // moving meshes, blocks [pblocks, gridDim) rotate the meshes into the next
  // step's geometry buffer (mv.dst) while the others compute this step
  int variant, pblocks, move_mesh;
  int step_from_ctrl;  // step = ctrl->step_base + step, ago_next = ctrl->ago_base + step
  MeshMoveArgs mv;
};

struct RebuildArgs {
  DevParams prm;
  int n, cap, W;
  int mixed;
  StateBuf cur, alt;
  AuxBuf acur, aalt;
  ListBuf oldl, newl;
  int have_old;
  int ncell;
  const int *cell_rank;  // linear cell -> Morton rank
  int *cell_count;       // [ncell]
  int *cell_start;       // [ncell+1]
  int *key;              // [cap] Morton rank of each particle's cell
  int *perm;             // [cap] new index -> old index
  double *xhold;
  int *scan_tmp;
  Ctrl *ctrl;
  // meshes
  int ntri;                    // all meshes
  TriGeom *tri;                // geometry of the rebuild step
  TriGeom *tri_alt;            // the other geometry buffer (nullptr if none)
  double *node_last;           // [ntri][9] nodes at the last rebuild
  int *tcount, *toff;          // per triangle: cells touched, offsets
  int *tpair_cell, *tpair_tri; // triangle-cell pairs
  int tpair_cap;
  int *tcell_count, *tcell_start, *tcell_list;
  int *cand, *ncand;
  int zero_vnode;              // setup: node velocities start at zero
};

// compute pair/gran/local: per particle the number of its half-list pairs
// (those whose partner has a higher tag), their offsets, and per pair the two
// tags and a record of ncol doubles: x_i[3], x_j[3], v_i[3], v_j[3], force on
// i[3], torque on i[3], periodic-image flag, r_i, r_j, history[hist_w]
constexpr int kPairLocalFixedCols = 21;
struct PairLocalArgs {
  int *count;      // [n]
  int *offset;     // [n + 1]
  int *tags;       // [npairs][2]
  double *rec;     // [npairs][ncol]
  int ncol;
};

// compute wall/gran/local: per particle the number of its mesh contact rows,
// their offsets, and per row the mesh, the triangle's index within its mesh
// and the particle's tag, and a record of ncol doubles: contact point[3],
// x_i[3], wall velocity at the contact[3], v_i[3], force on i[3], torque on
// i[3], unit normal from the particle centre to the contact point[3], the
// squared distance from the centre to the contact point, the radius,
// history[3]
constexpr int kWallLocalFixedCols = 23;
struct WallLocalArgs {
  int *count;      // [n]
  int *offset;     // [n + 1]
  int *ids;        // [nrows][3]
  double *rec;     // [nrows][ncol]
  int ncol;
};

// launchers (dem_kernels.hip)
hipError_t launch_rebuild(const RebuildArgs &a, hipStream_t s);
hipError_t launch_step(const StepArgs &a, hipStream_t s);
hipError_t launch_initial(const StepArgs &a, hipStream_t s);
hipError_t launch_pair_local_count(const StepArgs &a, const PairLocalArgs &o, hipStream_t s);
hipError_t launch_pair_local_fill(const StepArgs &a, const PairLocalArgs &o, hipStream_t s);
hipError_t launch_wall_local_count(const StepArgs &a, const WallLocalArgs &o, hipStream_t s);
hipError_t launch_wall_local_fill(const StepArgs &a, const WallLocalArgs &o, hipStream_t s);
// stop_at = v, run counters cleared (and the per-instance ones, nstat entries)
hipError_t launch_set_stop(Ctrl *ctrl, long long v, hipStream_t s,
                           unsigned long long *inst_stats = nullptr, int nstat = 0);
hipError_t launch_mesh_move(const MeshMoveArgs &a, hipStream_t s);
// step index and neighbor age read by graph-captured step launches
hipError_t launch_set_base(Ctrl *ctrl, long long step, long long ago, hipStream_t s);
// Mesh part of a rebuild, after the particle gather and build. Phase 1:
// element geometry refreshed from the nodes, node positions stored, and the
// number of (triangle, cell) pairs counted; the total lands in *total (device).
hipError_t launch_mesh_bin_count(const RebuildArgs &a, int *total, hipStream_t s);
// Phase 2 (tpair_cap >= total): triangle lists per cell, candidate triangles
// per particle, slot compaction and removal of slots whose triangle is no
// longer a candidate.
hipError_t launch_mesh_bin_fill(const RebuildArgs &a, hipStream_t s);
int scan_tmp_size(int ncell);
// out[0..n) = exclusive prefix sums of in[0..n), out[n] = total; tmp holds
// scan_tmp_size(n) ints
hipError_t launch_exclusive_scan(const int *in, int *out, int *tmp, int n, hipStream_t s);

}  // namespace dem
