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
constexpr uint32_t kFlagBit = 0x80000000u;  // contact flag stored in the neighbor entry
constexpr uint32_t kIndexMask = 0x7fffffffu;
constexpr int kMetaCountMask = 0xffff;       // meta[i]: low 16 bits = list length
constexpr int kMetaWallShift = 16;           // bits 16..21: wall history present

// Scalar run parameters, passed by value to every kernel.
struct DevParams {
  double dt, dtf, dtfrotate;
  double triggersq;        // 0.25 * skin * skin (neighbor.cpp:290)
  double skin, cdf;
  double lo[3], L[3], halfL[3];
  int per[3];
  int rolling, tdamp, limitf, torsion;
  int has_grav, grav_first;
  double gacc[3];
  int nwalls;
  int wall_axis[kMaxWalls];
  double wall_pos[kMaxWalls];
  int wall_type[kMaxWalls];
  int wall_model[kMaxWalls];  // bit 0 rolling cdt, 1 tangential damping, 2 limitForce, 3 torsion
  double cutneighmax;
  int ntp1;                // ntypes + 1 (row length of the material tables)
  // cell grid
  int nc[3];
  double cinv[3];
};

// Mixed material tables, [ntp1*ntp1] each, FP64 and FP32 copies.
struct MatTables {
  double *Yeff, *Geff, *beta, *mu, *mur;
  float *fYeff, *fGeff, *fbeta, *fmu, *fmur;
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
  int pad;
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
  void *whist;    // wall shear history, [wall][3][cap], double or float
};

struct ListBuf {
  uint32_t *nbr;  // [W][cap] neighbor index | contact flag
  void *hist;     // [W][3][cap] tangential history, double or float
};

struct StepArgs {
  DevParams prm;
  MatTables mat;
  int n, cap, W;
  int mixed;
  StateBuf in, out;
  AuxBuf aux;
  ListBuf list;
  double *xhold;     // [3][cap]
  double *fout;      // [6][cap] f and torque when store_f
  Ctrl *ctrl;
  long long step;
  int shearupdate, do_final, do_initial, store_f, count;
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
};

// launchers (dem_kernels.hip)
hipError_t launch_rebuild(const RebuildArgs &a, hipStream_t s);
hipError_t launch_step(const StepArgs &a, hipStream_t s);
hipError_t launch_initial(const StepArgs &a, hipStream_t s);
hipError_t launch_set_stop(Ctrl *ctrl, long long v, hipStream_t s);
int scan_tmp_size(int ncell);

}  // namespace dem
