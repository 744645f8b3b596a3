// Superquadric DEM step on the GPU: device data and launchers
// (sqdem_kernels.hip). FP64 throughout.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sqdk {

constexpr int kHist = 10;   // state, first OBB inequality, contact point[3], alpha_i, alpha_j, shear[3]
constexpr int kRec = 9;     // per pair: force on i, torque on i, torque on j
constexpr int kMaxPlanes = 6;
constexpr int kNumStats = 8;
// per-step counters
enum {
  kStatPairTouch = 0,   // touching pairs
  kStatWallTouch = 1,   // touching particle-plane contacts
  kStatNewton = 2,      // Newton iterations
  kStatChecked = 3,     // pairs whose bounding spheres overlap (contact detection ran)
  kStatFresh = 4,       // solves of pairs far apart last step that passed the OBB test
  kStatFreshIters = 5,  // their Newton iterations
  kStatFallback = 6,    // network starts replaced by the homotopy
  kStatRebuild = 7      // set by the displacement check
};

struct Dev {
  int n;
  double *x, *v, *w, *L, *q, *shape, *blk, *mass, *rv, *rb, *I, *f, *t, *xhold, *wrap;
  const int *type;
  double lo[3], hi[3], len[3];
  int per[3];
  double dt, skin;
  int tdamp, limit;
  int t1;  // ntypes + 1 (material tables are [t1][t1])
  const double *Y, *G, *B, *mu;
  int nplanes, pax[kMaxPlanes], ptype[kMaxPlanes];
  double ppos[kMaxPlanes];
  int has_grav, grav_first;
  double g[3];
  // half list (i < j), CSR by i, j ascending
  int64_t np;
  const int *first, *pi, *pj;
  int *flag;
  double *hist, *rec;
  // the same pairs listed by j, pair index ascending
  const int *rfirst, *rlist;
  double *whist;  // [n][nplanes][3]
  unsigned long long *stats;
  // network start points of pairs far apart last step (or null)
  const double *init;
  const int *fresh_idx;
  int max_iter;
};

struct Grid {
  int nc[3];
  double cs[3];
  int ncell;
  int *cell_of, *count, *start, *fill, *list;
  int *scan_tmp;
};

struct Build {
  // new list
  int *count, *first, *pi, *pj, *flag;
  double *hist;
  int cap;
  // previous list (history carried for flagged pairs)
  const int *ofirst, *opj, *oflag;
  const double *ohist;
  int onp;
  int *rcount, *rfirst, *rfill, *rlist;
  int *overflow;
};

hipError_t launch_initial(const Dev &d, int *flag, hipStream_t s);
hipError_t launch_wrap(const Dev &d, hipStream_t s);
// bins particles into the grid (cells sorted by particle index)
hipError_t launch_bin(const Dev &d, const Grid &g, hipStream_t s);
// counts each particle's partners j > i within the cutoff into b.count
hipError_t launch_pair_count(const Dev &d, const Grid &g, const Build &b, hipStream_t s);
// fills the list (b.first scanned), sorts each row, carries history, builds the by-j lists
hipError_t launch_pair_fill(const Dev &d, const Grid &g, const Build &b, hipStream_t s);
hipError_t launch_reverse(const Dev &d, const Build &b, int *scan_tmp, hipStream_t s);
hipError_t launch_zero_stats(unsigned long long *stats, hipStream_t s);
// totals[k] += stats[k]
hipError_t launch_accumulate(const unsigned long long *stats, unsigned long long *totals, hipStream_t s);
// pairs far apart last step whose OBBs overlap: compact list and their poses
hipError_t launch_fresh_scan(const Dev &d, int *fresh_idx, int *fresh_count, double *pose_a, double *pose_b,
                             hipStream_t s);
hipError_t launch_pairs(const Dev &d, int shearupdate, hipStream_t s);
// sums pair records, adds plane walls and gravity, and (integrate) the final half step
hipError_t launch_gather(const Dev &d, int shearupdate, int integrate, hipStream_t s);

}  // namespace sqdk
