// Host side of the DEM step library: the C ABI of include/liggghts_amd.h.
//
// A handle owns device buffers for one granular system and drives the
// per-timestep sequence of Verlet::run (verlet.cpp:264-392):
//   setup : carry history out, wrap, Morton sort, neighbor build, force pass
//   steps : fused [pair + walls + gravity + final integrate + next initial
//           integrate + displacement check] kernel, one launch per step.
// Steps are enqueued ahead in batches; a kernel whose step index reaches the
// control block's stop_at (raised by the previous step's displacement check)
// returns at once, and the host rebuilds before re-launching that step. The
// host therefore never waits on the GPU between two steps that need no rebuild.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/liggghts_amd.h"
#include "dem_device.h"
#include "mesh_host.h"

using dem::AuxBuf;
using dem::Ctrl;
using dem::ListBuf;
using dem::StateBuf;

namespace {

constexpr long long kNever = 0x3fffffffffffffffLL;
constexpr int64_t kMotionSteps = 65536;  // steps per run piece with time-dependent movers
constexpr double kMyPi = 3.14159265358979323846;  // math_const.h

struct Wall {
  int axis;
  double pos;
  int type;
  int model;  // contact-model bits (dem::kModel*)
};

int model_bits(const dem_contact_model &m) {
  return (m.rolling_cdt ? 1 : 0) | (m.tangential_damping ? 2 : 0) | (m.limit_force ? 4 : 0) |
         (m.torsion_torque ? 8 : 0) | ((m.normal_model & 3) << dem::kModelNormalShift) |
         (m.tangential_model ? dem::kModelNoHistory : 0) | (m.cohesion_model ? dem::kModelSjkr : 0) |
         (m.kt_to_kn ? dem::kModelKtToKn : 0) | (m.absolute_damping ? dem::kModelAbsDamping : 0) |
         (m.tangential_reduce ? dem::kModelTangReduce : 0) | (m.rolling_epsd2 ? dem::kModelEpsd2 : 0);
}

const char *normal_model_name(int nm) {
  static const char *names[4] = {"hertz", "hooke", "hertz/stiffness", "hooke/stiffness"};
  return names[nm & 3];
}

std::string g_create_error;

}  // namespace

struct dem_handle {
  dem_params p{};
  std::string err;
  hipStream_t stream = nullptr;
  bool mixed = false;
  size_t fsz = 8;  // bytes of one history / omega component
  int hw = 3;      // history values per pair / plane contact (6 with epsd2)

  // materials
  int ntypes = 0;
  bool mat_set = false;
  // ninst blocks of (ntypes+1)^2 each, in dem::MatTables order
  std::vector<double> hYeff, hGeff, hbeta, hmu, hmur, hkn, hkt, hgn, hgt, hced, hloge;
  bool mat_dirty = false;     // host tables newer than the device copy
  bool have_elastic = false;  // youngsModulus / poissonsRatio / coefficientRestitution given
  bool have_stiff = false;    // kn / kt / gamman / gammat given
  bool have_ced = false;      // cohesionEnergyDensity given
  double char_vel = 0.0;      // characteristicVelocity (0: not given)
  dem_contact_model pair_model{};  // pair_style gran model switches
  // independent simulations batched in this handle
  int ninst = 1;
  std::vector<int64_t> inst_count;  // particles of each instance, in caller order
  int ncell_local = 0;              // cells of one instance's grid
  unsigned long long *d_inst_stats = nullptr;  // [2][ninst]
  dem::MatTables mat{};
  void *mat_block = nullptr;

  std::vector<Wall> walls;

  // triangle meshes (host copies; device copies live with the particles)
  std::vector<dem::HostMesh> meshes;
  int ntri = 0;
  // Triangle geometry, two buffers when a mesh moves: a step reads one and
  // its trailing blocks write the next step's geometry into the other.
  // tri_cur holds the geometry of the current state; without moving meshes
  // both entries point to one buffer.
  dem::TriGeom *d_tri_buf[2] = {nullptr, nullptr};
  int tri_cur = 0;
  bool pending_mesh_rebuild = false;  // a moving mesh asked for a rebuild at the next step
  double *d_node_last = nullptr;
  // time-dependent movers (wiggle, viblin, vibrot): their motion at every
  // step of the current run (dem::kMotionStride values per mesh), and the
  // FixMoveMesh time so far (time_ += dt each step over the fix's lifetime)
  double *d_motion = nullptr;
  int64_t motion_cap = 0;  // run steps the device table holds
  double mesh_time = 0.0;
  int *d_copl_off = nullptr, *d_copl_idx = nullptr;
  int *d_tcount = nullptr, *d_toff = nullptr;
  int *d_tpair_cell = nullptr, *d_tpair_tri = nullptr, *d_tcell_list = nullptr;
  int tpair_cap = 0;
  int *d_tcell_count = nullptr, *d_tcell_start = nullptr;
  int *d_cand = nullptr, *d_ncand = nullptr;
  int *d_total = nullptr;
  int *h_total = nullptr;  // pinned
  bool zero_vnode = false;

  // particles
  int64_t n = 0;
  int cap = 0;
  int W = 32;
  StateBuf st[2]{};
  AuxBuf aux[2]{};
  ListBuf lst[2]{};
  int cur_st = 0, cur_aux = 0, cur_l = 0;
  bool have_list = false;
  bool setup_done = false;
  bool can_continue = false;  // fout holds the last force pass in the current order
  double *xhold = nullptr, *fout = nullptr;
  double rmax = 0.0;
  double rmax_future = 0.0;  // dem_set_max_radius: particles still to come
  // particles waiting to join at the next run's first step (dem_insert_particles)
  struct Pending {
    std::vector<double> x, v, w, r, m;
    std::vector<int> type, tag;
  } pend;

  // cell grid
  int nc[3] = {0, 0, 0};
  int ncell = 0;
  double cutneighmax = 0.0;
  int *cell_rank = nullptr, *cell_count = nullptr, *cell_start = nullptr;
  int *key = nullptr, *perm = nullptr, *scan_tmp = nullptr;

  Ctrl *ctrl = nullptr;
  Ctrl *ctrl_host = nullptr;

  // counters
  int64_t ntimestep = 0, nrebuilds = 0;
  int64_t last_build = 0;  // timestep of the last neighbor build
  int64_t npairs = 0, ncontacts = 0, nwall_contacts = 0, nmesh_cand = 0;
  int maxcount = 0;
  bool timing = false;
  // tile shape of the step kernel: variant (dem::kStepVariant*) and a
  // persistent grid of `persist` blocks per compute unit (0: one block per tile)
  int variant = 0, persist = 0, ncu = 256;
  // step batches captured as HIP graphs, one per buffer parity and shape
  bool use_graphs = false;
  struct GraphEntry {
    int key;
    hipGraphExec_t exec;
  };
  std::vector<GraphEntry> graphs;
  struct Tuner {
    bool active = false;
    std::vector<std::pair<int, int>> cand;  // (variant, persist)
    std::vector<double> us;                 // fastest measured microseconds per step
    int idx = 0, tries = 0;
    int round = 0;                          // passes over the candidates so far
  } tuner;
  double kernel_ms = 0.0;
  int64_t kernel_launches = 0;
  double rebuild_ms = 0.0;
  std::vector<hipEvent_t> ev;

  std::vector<void *> allocs;
};

namespace {

int fail(dem_handle *h, int code, const std::string &msg) {
  if (h)
    h->err = msg;
  else
    g_create_error = msg;
  return code;
}

#define HIPCHK(h, call)                                                            \
  do {                                                                             \
    hipError_t e_ = (call);                                                        \
    if (e_ != hipSuccess)                                                          \
      return fail((h), DEM_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

int dalloc(dem_handle *h, void **ptr, size_t bytes) {
  *ptr = nullptr;
  if (bytes == 0) bytes = 256;
  hipError_t e = hipMalloc(ptr, bytes);
  if (e != hipSuccess)
    return fail(h, DEM_ERR_HIP, std::string("hipMalloc failed: ") + hipGetErrorString(e));
  h->allocs.push_back(*ptr);
  return DEM_OK;
}

void clear_graphs(dem_handle *h) {
  for (auto &g : h->graphs) hipGraphExecDestroy(g.exec);
  h->graphs.clear();
}

// frees one allocation made by dalloc
void dfree(dem_handle *h, void *p) {
  if (!p) return;
  auto it = std::find(h->allocs.begin(), h->allocs.end(), p);
  if (it != h->allocs.end()) h->allocs.erase(it);
  hipFree(p);
}


void free_particles(dem_handle *h) {
  clear_graphs(h);
  // everything except the material block and control block
  std::vector<void *> keep;
  for (void *p : h->allocs) {
    if (p == h->mat_block || p == (void *)h->ctrl) {
      keep.push_back(p);
      continue;
    }
    hipFree(p);
  }
  h->allocs = keep;
  for (int b = 0; b < 2; b++) {
    h->st[b] = StateBuf{};
    h->aux[b] = AuxBuf{};
    h->lst[b] = ListBuf{};
  }
  h->xhold = h->fout = nullptr;
  h->d_tri_buf[0] = h->d_tri_buf[1] = nullptr;
  h->d_node_last = nullptr;
  h->d_motion = nullptr;
  h->motion_cap = 0;
  h->d_copl_off = h->d_copl_idx = nullptr;
  h->d_tcount = h->d_toff = h->d_tpair_cell = h->d_tpair_tri = h->d_tcell_list = nullptr;
  h->tpair_cap = 0;
  h->d_tcell_count = h->d_tcell_start = h->d_cand = h->d_ncand = h->d_total = nullptr;
  h->cell_rank = h->cell_count = h->cell_start = nullptr;
  h->key = h->perm = h->scan_tmp = nullptr;
  h->ncell = 0;
  h->n = 0;
  h->cap = 0;
  h->have_list = false;
  h->setup_done = false;
  h->can_continue = false;
}

// any pair, wall or mesh contact model beyond Hertz / tangential history
bool uses_generic_models(const dem_handle *h) {
  int bits = model_bits(h->pair_model);
  for (const Wall &w : h->walls) bits |= w.model;
  for (const auto &m : h->meshes) bits |= m.model;
  return (bits & dem::kModelGenericMask) != 0;
}

// the host material tables in dem::MatTables order
std::array<std::vector<double> *, dem::kNumMatTables> host_tables(dem_handle *h) {
  return {&h->hYeff, &h->hGeff, &h->hbeta, &h->hmu, &h->hmur, &h->hkn,
          &h->hkt,   &h->hgn,   &h->hgt,   &h->hced, &h->hloge};
}

dem::DevParams make_devparams(const dem_handle *h) {
  dem::DevParams d{};
  const dem_params &p = h->p;
  d.dt = p.dt;
  d.dtf = 0.5 * p.dt * 1.0;  // fix_nve.cpp:86 (ftm2v = 1 in SI)
  d.dtfrotate = d.dtf / 0.4; // fix_nve_sphere.cpp:160
  d.triggersq = 0.25 * p.skin * p.skin;
  d.neigh_every = p.neigh_every > 0 ? p.neigh_every : 1;
  d.neigh_check = p.neigh_no_check ? 0 : 1;
  d.skin = p.skin;
  d.cdf = p.cdf;
  d.cdm = p.cdf * p.cdf;
  d.close_band = p.cdf > 1.0 ? 1 : 0;
  for (int k = 0; k < 3; k++) {
    d.lo[k] = p.boxlo[k];
    d.L[k] = p.boxhi[k] - p.boxlo[k];
    d.halfL[k] = 0.5 * d.L[k];
    d.per[k] = p.periodic[k] ? 1 : 0;
    if (!p.periodic[k])
      for (int side = 0; side < 2; side++)
        if (!((p.shrink_faces >> (2 * k + side)) & 1)) d.lost_faces |= 1 << (2 * k + side);
    d.nc[k] = h->nc[k];
    d.cinv[k] = h->nc[k] > 0 ? h->nc[k] / d.L[k] : 0.0;
  }
  d.pair_model = model_bits(h->pair_model);
  d.char_vel = h->char_vel;
  d.generic = (uses_generic_models(h) || h->hw != 3) ? 1 : 0;
  d.hist_w = h->hw;
  d.has_grav = p.has_gravity;
  d.grav_first = p.gravity_first;
  for (int k = 0; k < 3; k++) d.gacc[k] = p.grav_acc[k];
  d.nwalls = (int)h->walls.size();
  for (size_t w = 0; w < h->walls.size(); w++) {
    d.wall_axis[w] = h->walls[w].axis;
    d.wall_pos[w] = h->walls[w].pos;
    d.wall_type[w] = h->walls[w].type;
    d.wall_model[w] = h->walls[w].model;
  }
  d.cutneighmax = h->cutneighmax;
  d.ntp1 = h->ntypes + 1;
  d.mat_stride = d.ntp1 * d.ntp1;
  d.ninst = h->ninst;
  d.ncell_local = h->ncell_local;
  d.nmesh = (int)h->meshes.size();
  int t0 = 0;
  for (size_t m = 0; m < h->meshes.size(); m++) {
    d.mesh_tri0[m] = t0;
    d.mesh_ntri[m] = h->meshes[m].ntri;
    d.mesh_wtype[m] = h->meshes[m].wall_type;
    d.mesh_model[m] = h->meshes[m].model;
    d.mesh_moving[m] = h->meshes[m].moving ? 1 : 0;
    if (h->meshes[m].stress) d.mesh_stress |= 1 << m;
    for (int k = 0; k < 3; k++) d.mesh_pref[m][k] = h->meshes[m].p_ref[k];
    t0 += h->meshes[m].ntri;
  }
  d.mesh_precision = h->meshes.empty() ? 1e-8 : h->meshes[0].precision;
  return d;
}

uint32_t spread_bits(uint32_t v) {
  // 10-bit value -> every third bit (Morton interleave)
  v &= 0x3ff;
  v = (v | (v << 16)) & 0x030000FF;
  v = (v | (v << 8)) & 0x0300F00F;
  v = (v | (v << 4)) & 0x030C30C3;
  v = (v | (v << 2)) & 0x09249249;
  return v;
}

// Cell grid with cells no smaller than cutneighmax, so the 27-cell stencil
// covers every pair within (r_i + r_j) * cdf + skin. Cells are ranked in
// Morton order; particles are counting-sorted by that rank at each rebuild.
int setup_grid(dem_handle *h) {
  // pair_gran.cpp:623-635, with future particles counted as dynamic (:538-551)
  const double rm = std::max(h->rmax, h->rmax_future);
  h->cutneighmax = (rm + rm) * h->p.cdf + h->p.skin;
  long long total = 1;
  for (int d = 0; d < 3; d++) {
    const double L = h->p.boxhi[d] - h->p.boxlo[d];
    int nc = (int)(L / h->cutneighmax);
    if (nc < 1) nc = 1;
    if (h->p.periodic[d] && nc < 3) nc = 1;
    if (nc > 1024) nc = 1024;
    if (h->p.periodic[d] && nc > 1 && L < 2.0 * h->cutneighmax)
      return fail(h, DEM_ERR_ARG, "periodic box shorter than twice the neighbor cutoff");
    h->nc[d] = nc;
    total *= nc;
  }
  if (total * h->ninst > (1LL << 27)) return fail(h, DEM_ERR_ARG, "cell grid too large");
  const int ncell = (int)total;
  std::vector<std::pair<uint32_t, int>> code(ncell);
  for (int cz = 0; cz < h->nc[2]; cz++)
    for (int cy = 0; cy < h->nc[1]; cy++)
      for (int cx = 0; cx < h->nc[0]; cx++) {
        const int lin = cx + h->nc[0] * (cy + h->nc[1] * cz);
        code[lin] = {spread_bits(cx) | (spread_bits(cy) << 1) | (spread_bits(cz) << 2), lin};
      }
  std::sort(code.begin(), code.end());
  std::vector<int> rank(ncell);
  for (int r = 0; r < ncell; r++) rank[code[r].second] = r;
  // every instance has a copy of this grid: the rank table is per instance,
  // the counts and starts cover all instances
  const int nglobal = ncell * h->ninst;
  int rc;
  if ((rc = dalloc(h, (void **)&h->cell_rank, sizeof(int) * ncell))) return rc;
  if ((rc = dalloc(h, (void **)&h->cell_count, sizeof(int) * nglobal))) return rc;
  if ((rc = dalloc(h, (void **)&h->cell_start, sizeof(int) * (nglobal + 1)))) return rc;
  if ((rc = dalloc(h, (void **)&h->scan_tmp,
                   sizeof(int) * dem::scan_tmp_size(std::max(nglobal, h->ntri + 1)))))
    return rc;
  HIPCHK(h, hipMemcpy(h->cell_rank, rank.data(), sizeof(int) * ncell, hipMemcpyHostToDevice));
  h->ncell_local = ncell;
  h->ncell = nglobal;
  return DEM_OK;
}

dem::RebuildArgs rebuild_args(dem_handle *h) {
  dem::RebuildArgs a{};
  a.prm = make_devparams(h);
  a.n = (int)h->n;
  a.cap = h->cap;
  a.W = h->W;
  a.mixed = h->mixed ? 1 : 0;
  a.cur = h->st[h->cur_st];
  a.alt = h->st[1 - h->cur_st];
  a.acur = h->aux[h->cur_aux];
  a.aalt = h->aux[1 - h->cur_aux];
  a.oldl = h->lst[h->cur_l];
  a.newl = h->lst[1 - h->cur_l];
  a.have_old = h->have_list ? 1 : 0;
  a.ncell = h->ncell;
  a.cell_rank = h->cell_rank;
  a.cell_count = h->cell_count;
  a.cell_start = h->cell_start;
  a.key = h->key;
  a.perm = h->perm;
  a.xhold = h->xhold;
  a.scan_tmp = h->scan_tmp;
  a.ctrl = h->ctrl;
  a.ntri = h->ntri;
  a.tri = h->d_tri_buf[h->tri_cur];
  a.tri_alt = h->d_tri_buf[1 - h->tri_cur] != a.tri ? h->d_tri_buf[1 - h->tri_cur] : nullptr;
  a.node_last = h->d_node_last;
  a.tcount = h->d_tcount;
  a.toff = h->d_toff;
  a.tpair_cell = h->d_tpair_cell;
  a.tpair_tri = h->d_tpair_tri;
  a.tpair_cap = h->tpair_cap;
  a.tcell_count = h->d_tcell_count;
  a.tcell_start = h->d_tcell_start;
  a.tcell_list = h->d_tcell_list;
  a.cand = h->d_cand;
  a.ncand = h->d_ncand;
  a.zero_vnode = h->zero_vnode ? 1 : 0;
  return a;
}

dem::MeshMoveArgs mesh_move_args(const dem_handle *h, long long step, int guard, int decide);

// in_buf: particle state buffer read by the step; tri_buf: geometry buffer of
// the step (the next step's geometry goes to the other one)
dem::StepArgs step_args(dem_handle *h, int in_buf, int tri_buf) {
  dem::StepArgs a{};
  a.prm = make_devparams(h);
  a.mat = h->mat;
  a.n = (int)h->n;
  a.cap = h->cap;
  a.W = h->W;
  a.mixed = h->mixed ? 1 : 0;
  a.in = h->st[in_buf];
  a.out = h->st[1 - in_buf];
  a.aux = h->aux[h->cur_aux];
  a.list = h->lst[h->cur_l];
  a.mesh.tri = h->d_tri_buf[tri_buf];
  a.mesh.copl_off = h->d_copl_off;
  a.mesh.copl_idx = h->d_copl_idx;
  a.mesh.cand = h->d_cand;
  a.mesh.ncand = h->d_ncand;
  a.xhold = h->xhold;
  a.fout = h->fout;
  a.ctrl = h->ctrl;
  a.inst_stats = h->d_inst_stats;
  // the sibling models run in shapes 0 and 5 only
  a.variant = (a.prm.generic && h->variant != 5) ? 0 : h->variant;
  {
    const int B = dem::kStepVariantBlock[a.variant];
    const int ntiles = (int)((std::max<int64_t>(h->n, 1) + B - 1) / B);
    a.pblocks = h->persist > 0 ? std::min(ntiles, h->persist * h->ncu) : ntiles;
  }
  a.move_mesh = 0;
  if (h->ntri > 0) {
    a.mv = mesh_move_args(h, 0, 0, 1);
    a.mv.tri = h->d_tri_buf[tri_buf];
    a.mv.dst = h->d_tri_buf[1 - tri_buf];
  }
  return a;
}

bool any_moving_mesh(const dem_handle *h) {
  for (const auto &m : h->meshes)
    if (m.moving) return true;
  return false;
}

bool has_timed_movers(const dem_handle *h) {
  for (const auto &m : h->meshes)
    if (m.mover == dem::kMoverWiggle || m.mover == dem::kMoverVibLin || m.mover == dem::kMoverVibRot)
      return true;
  return false;
}

// The time-dependent movers' motion at run steps 1..nsteps, evaluated here
// (host libm, as the reference evaluates it): the mover's time t is
// accumulated as FixMoveMesh::initial_integrate does (time_ += dt,
// fix_move_mesh.cpp:226) and
//   wiggle  w cos(w t)                                  (mesh_mover_linear.cpp:203-218)
//   viblin  vA = sum_j -a_j w_j sin(w_j t + phi_j)       (:275-305)
//   vibrot  vR likewise; rotation by vR dt about the axis through the origin,
//           angular velocity axis vR                     (mesh_mover_rotation.cpp:397-430)
// Returns the time after the last step in *t_end.
int upload_mesh_motion(dem_handle *h, int64_t nsteps, double *t_end) {
  constexpr int S = dem::kMotionStride, M = dem::kMaxMeshes;
  std::vector<double> tab((size_t)(nsteps + 1) * M * S, 0.0);
  double T = h->mesh_time;
  for (int64_t k = 1; k <= nsteps; k++) {
    T = T + h->p.dt;
    for (size_t m = 0; m < h->meshes.size(); m++) {
      const dem::HostMesh &hm = h->meshes[m];
      double *e = &tab[((size_t)k * M + m) * S];
      if (hm.mover == dem::kMoverWiggle) {
        e[7] = hm.womega * std::cos(hm.womega * T);
      } else if (hm.mover == dem::kMoverVibLin || hm.mover == dem::kMoverVibRot) {
        double vA = 0.0;
        for (int j = 0; j < hm.vord; j++)
          vA = vA - hm.vampl[j] * hm.vomega[j] * std::sin(hm.vomega[j] * T + hm.vphi[j]);
        e[7] = vA;
        if (hm.mover == dem::kMoverVibRot) {
          // MultiNodeMesh::rotate(dAngle, axis, p), multi_node_mesh_I.h:620-641
          const double dAngle = vA * h->p.dt;
          double an[3] = {hm.axis[0], hm.axis[1], hm.axis[2]};
          const double sinv = 1. / std::sqrt(an[0] * an[0] + an[1] * an[1] + an[2] * an[2]);
          for (int q = 0; q < 3; q++) an[q] = sinv * an[q];
          e[0] = std::cos(dAngle * 0.5);
          for (int q = 0; q < 3; q++) e[1 + q] = an[q] * std::sin(dAngle * 0.5);
          for (int q = 0; q < 3; q++) e[4 + q] = hm.axis[q] * vA;
        }
      }
    }
  }
  *t_end = T;
  if (nsteps + 1 > h->motion_cap) {
    clear_graphs(h);  // captured launches hold the old table pointer
    if (h->d_motion) {
      hipFree(h->d_motion);
      h->allocs.erase(std::remove(h->allocs.begin(), h->allocs.end(), (void *)h->d_motion), h->allocs.end());
      h->d_motion = nullptr;
    }
    const int64_t cap = std::max<int64_t>(nsteps + 1, std::min<int64_t>(2 * h->motion_cap, kMotionSteps + 1));
    int rc = dalloc(h, (void **)&h->d_motion, sizeof(double) * (size_t)cap * M * S);
    if (rc) return rc;
    h->motion_cap = cap;
  }
  HIPCHK(h, hipStreamSynchronize(h->stream));  // the previous run's launches are done with it
  HIPCHK(h, hipMemcpy(h->d_motion, tab.data(), sizeof(double) * tab.size(), hipMemcpyHostToDevice));
  return DEM_OK;
}

// one step of fix move/mesh rotate for every moving mesh
dem::MeshMoveArgs mesh_move_args(const dem_handle *h, long long step, int guard, int decide) {
  dem::MeshMoveArgs a{};
  a.tri = h->d_tri_buf[h->tri_cur];
  a.dst = nullptr;
  a.node_last = h->d_node_last;
  a.ntri = h->ntri;
  a.triggersq = 0.25 * h->p.skin * h->p.skin;
  a.step = step;
  a.guard = guard;
  a.decide = decide;
  a.ctrl = h->ctrl;
  a.motion = h->d_motion;
  a.dt = h->p.dt;
  for (size_t m = 0; m < h->meshes.size(); m++) {
    const dem::HostMesh &hm = h->meshes[m];
    if (!hm.moving) continue;
    a.moving_mask |= 1 << m;
    a.kind[m] = hm.mover;
    if (hm.mover == dem::kMoverLinear || hm.mover == dem::kMoverWiggle || hm.mover == dem::kMoverVibLin) {
      for (int k = 0; k < 3; k++)
        a.tvec[m][k] = hm.mover == dem::kMoverLinear ? hm.vel[k]
                       : hm.mover == dem::kMoverWiggle ? hm.ampl[k] : hm.axis[k];
      continue;
    }
    if (hm.mover == dem::kMoverVibRot) {  // increments come from the motion table
      for (int k = 0; k < 3; k++) a.origin[m][k] = hm.origin[k];
      a.trans[m] = (hm.origin[0] * hm.origin[0] + hm.origin[1] * hm.origin[1] +
                    hm.origin[2] * hm.origin[2]) > 0.;
      continue;
    }
    // MultiNodeMesh::rotate(dAngle, axis, p), multi_node_mesh_I.h:620-641
    const double dAngle = hm.omega * h->p.dt;
    double axisNorm[3] = {hm.axis[0], hm.axis[1], hm.axis[2]};
    const double mg = std::sqrt(axisNorm[0] * axisNorm[0] + axisNorm[1] * axisNorm[1] +
                                axisNorm[2] * axisNorm[2]);
    const double sinv = 1. / mg;
    for (int k = 0; k < 3; k++) axisNorm[k] = sinv * axisNorm[k];
    a.dQ[m][0] = std::cos(dAngle * 0.5);
    for (int k = 0; k < 3; k++) a.dQ[m][k + 1] = axisNorm[k] * std::sin(dAngle * 0.5);
    for (int k = 0; k < 3; k++) {
      a.origin[m][k] = hm.origin[k];
      a.omegaVec[m][k] = hm.axis[k] * hm.omega;
    }
    a.trans[m] = (hm.origin[0] * hm.origin[0] + hm.origin[1] * hm.origin[1] +
                  hm.origin[2] * hm.origin[2]) > 0.;
  }
  return a;
}

int read_ctrl(dem_handle *h) {
  HIPCHK(h, hipMemcpyAsync(h->ctrl_host, h->ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return DEM_OK;
}

// Neighbor rebuild (verlet.cpp:313-336): FixContactHistory::pre_exchange,
// Domain::pbc, atom sort, Neighbor::build with history carry-in.
int rebuild(dem_handle *h) {
  hipEvent_t e0, e1;
  if (h->timing) {
    HIPCHK(h, hipEventCreate(&e0));
    HIPCHK(h, hipEventCreate(&e1));
    HIPCHK(h, hipEventRecord(e0, h->stream));
  }
  dem::RebuildArgs a = rebuild_args(h);
  HIPCHK(h, dem::launch_rebuild(a, h->stream));
  if (h->ntri > 0) {
    HIPCHK(h, dem::launch_mesh_bin_count(a, h->d_total, h->stream));
    HIPCHK(h, hipMemcpyAsync(h->h_total, h->d_total, sizeof(int), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    const int total = *h->h_total;
    if (total < 0 || (int64_t)total * h->ninst > 0x7fffffffLL)
      return fail(h, DEM_ERR_STATE, "mesh binning overflow");
    if (total > h->tpair_cap) {
      // grow the (triangle, cell) pair arrays; the old ones stay allocated until
      // the particle arrays are freed
      const int cap = total + total / 4 + 1024;
      int rc2;
      if ((rc2 = dalloc(h, (void **)&h->d_tpair_cell, sizeof(int) * cap))) return rc2;
      if ((rc2 = dalloc(h, (void **)&h->d_tpair_tri, sizeof(int) * cap))) return rc2;
      // every instance gets its own copy of the pairs in its cells
      if ((rc2 = dalloc(h, (void **)&h->d_tcell_list, sizeof(int) * (size_t)cap * h->ninst)))
        return rc2;
      h->tpair_cap = cap;
      a.tpair_cell = h->d_tpair_cell;
      a.tpair_tri = h->d_tpair_tri;
      a.tcell_list = h->d_tcell_list;
    }
    a.tpair_cap = total;  // the fill phase takes the exact pair count here
    HIPCHK(h, dem::launch_mesh_bin_fill(a, h->stream));
  }
  if (h->timing) HIPCHK(h, hipEventRecord(e1, h->stream));
  int rc = read_ctrl(h);
  if (rc) return rc;
  if (h->ctrl_host->mesh_overflow > 0) {
    char buf[200];
    snprintf(buf, sizeof buf, "a particle has %d candidate mesh triangles (limit %d)",
             h->ctrl_host->mesh_overflow, dem::kMaxMeshCand);
    return fail(h, DEM_ERR_OVERFLOW, buf);
  }
  if (h->timing) {
    float ms = 0.f;
    HIPCHK(h, hipEventElapsedTime(&ms, e0, e1));
    h->rebuild_ms += ms;
    hipEventDestroy(e0);
    hipEventDestroy(e1);
  }
  if (h->ctrl_host->nlost > 0) {
    // Thermo::lost_check with thermo_modify lost error (thermo.cpp:434-441)
    char buf[160];
    snprintf(buf, sizeof buf, "Lost atoms: original %lld current %lld", (long long)h->n,
             (long long)(h->n - (int64_t)h->ctrl_host->nlost));
    return fail(h, DEM_ERR_STATE, buf);
  }
  if (h->ctrl_host->overflow > 0) {
    char buf[200];
    snprintf(buf, sizeof buf,
             "Neighbor list overflow, boost neigh_modify one (a particle has %d neighbors, "
             "limit %d)",
             h->ctrl_host->overflow, h->W);
    return fail(h, DEM_ERR_OVERFLOW, buf);
  }
  h->cur_st = 1 - h->cur_st;
  h->cur_aux = 1 - h->cur_aux;
  h->cur_l = 1 - h->cur_l;
  h->have_list = true;
  h->nrebuilds++;
  h->last_build = h->ntimestep;
  h->maxcount = h->ctrl_host->maxcount;
  h->npairs = (int64_t)(h->ctrl_host->npairs_full / 2);
  h->nmesh_cand = (int64_t)h->ctrl_host->nmesh_cand;
  return DEM_OK;
}

int force_pass(dem_handle *h, int shearupdate) {
  HIPCHK(h, dem::launch_set_stop(h->ctrl, kNever, h->stream, h->d_inst_stats, 2 * h->ninst));
  dem::StepArgs a = step_args(h, h->cur_st, h->tri_cur);
  a.step = 0;
  a.shearupdate = shearupdate;
  a.do_final = 0;
  a.do_initial = 0;
  a.store_f = 1;
  a.count = 1;
  HIPCHK(h, dem::launch_step(a, h->stream));
  int rc = read_ctrl(h);
  if (rc) return rc;
  h->ncontacts = (int64_t)(h->ctrl_host->ncontact_full / 2);
  h->nwall_contacts = (int64_t)h->ctrl_host->nwall;
  return DEM_OK;
}

// Y, nu and e may be absent together (models that do not read them: the
// stiffness models take kn, kt, gamman, gammat instead)
int check_material_args(dem_handle *h, int ntypes, const double *Y, const double *nu,
                        const double *e, const double *mu) {
  if (ntypes < 1 || ntypes > dem::kMaxTypes || !mu)
    return fail(h, DEM_ERR_ARG, "bad material arguments");
  if (!Y && !nu && !e) return DEM_OK;
  if (!Y || !nu || !e)
    return fail(h, DEM_ERR_ARG, "youngsModulus, poissonsRatio and coefficientRestitution go together");
  // range checks of global_properties.cpp:360-372, :393-396, :316-320
  for (int i = 0; i < ntypes; i++) {
    if (Y[i] < 5e6)
      return fail(h, DEM_ERR_ARG, "youngsModulus >= 5e6 required for SI units");
    if (Y[i] > 1e9)
      return fail(h, DEM_ERR_ARG, "youngsModulus <= 1e9 required for SI units");
    if (nu[i] < 0. || nu[i] > 0.5) return fail(h, DEM_ERR_ARG, "0 <= poissonsRatio <= 0.5 required");
  }
  for (int k = 0; k < ntypes * ntypes; k++)
    if (e[k] <= 0.05 || e[k] > 1)
      return fail(h, DEM_ERR_ARG, "0.05 < coefficientRestitution <= 1 required");
  return DEM_OK;
}

// mixes one material block (global_properties.cpp:426-433, :455-464, :518-541)
void mix_block(dem_handle *h, int block, const double *Y, const double *nu, const double *e,
               const double *mu, const double *mur) {
  const int ntypes = h->ntypes, t1 = ntypes + 1;
  const size_t b0 = (size_t)block * t1 * t1;
  for (int i = 1; i <= ntypes; i++)
    for (int j = 1; j <= ntypes; j++) {
      const size_t k = b0 + i * t1 + j;
      if (Y) {
        const double Yi = Y[i - 1], Yj = Y[j - 1], vi = nu[i - 1], vj = nu[j - 1];
        h->hYeff[k] = 1. / ((1. - std::pow(vi, 2.)) / Yi + (1. - std::pow(vj, 2.)) / Yj);
        h->hGeff[k] = 1. / (2. * (2. - vi) * (1. + vi) / Yi + 2. * (2. - vj) * (1. + vj) / Yj);
        const double le = std::log(e[(i - 1) * ntypes + (j - 1)]);
        h->hbeta[k] = le / std::sqrt(std::pow(le, 2.) + std::pow(kMyPi, 2.));
        h->hloge[k] = le;  // createCoeffRestLog (global_properties.cpp:502-520)
      }
      h->hmu[k] = mu[(i - 1) * ntypes + (j - 1)];
      h->hmur[k] = mur ? mur[(i - 1) * ntypes + (j - 1)] : 0.0;
    }
}

// device copy of the material tables (FP64 and FP32), all instance blocks
int upload_materials(dem_handle *h) {
  if (!h->mat_dirty) return DEM_OK;
  clear_graphs(h);  // captured launches hold the old table pointers
  if (h->mat_block) {
    hipFree(h->mat_block);
    h->allocs.erase(std::remove(h->allocs.begin(), h->allocs.end(), h->mat_block), h->allocs.end());
    h->mat_block = nullptr;
  }
  constexpr int NT = dem::kNumMatTables;
  const size_t nn = h->hYeff.size();
  std::vector<char> blob(nn * NT * (8 + 4));
  double *dd = (double *)blob.data();
  float *ff = (float *)(blob.data() + nn * NT * 8);
  const auto src = host_tables(h);
  for (int m = 0; m < NT; m++)
    for (size_t k = 0; k < nn; k++) {
      dd[m * nn + k] = (*src[m])[k];
      ff[m * nn + k] = (float)(*src[m])[k];
    }
  int rc = dalloc(h, &h->mat_block, blob.size());
  if (rc) return rc;
  h->mat_block = h->allocs.back();
  HIPCHK(h, hipMemcpy(h->mat_block, blob.data(), blob.size(), hipMemcpyHostToDevice));
  double *db = (double *)h->mat_block;
  float *fb = (float *)((char *)h->mat_block + nn * NT * 8);
  double **dptr[NT] = {&h->mat.Yeff, &h->mat.Geff, &h->mat.beta, &h->mat.mu,  &h->mat.mur, &h->mat.kn,
                       &h->mat.kt,   &h->mat.gn,   &h->mat.gt,   &h->mat.ced, &h->mat.loge};
  float **fptr[NT] = {&h->mat.fYeff, &h->mat.fGeff, &h->mat.fbeta, &h->mat.fmu,  &h->mat.fmur, &h->mat.fkn,
                      &h->mat.fkt,   &h->mat.fgn,   &h->mat.fgt,   &h->mat.fced, &h->mat.floge};
  for (int m = 0; m < NT; m++) {
    *dptr[m] = db + m * nn;
    *fptr[m] = fb + m * nn;
  }
  h->mat_dirty = false;
  return DEM_OK;
}

// resizes the host tables to ninst blocks, repeating block 0
void replicate_material_blocks(dem_handle *h) {
  const size_t bs = (size_t)(h->ntypes + 1) * (h->ntypes + 1);
  for (auto *t : host_tables(h)) {
    t->resize(bs * h->ninst);
    for (int q = 1; q < h->ninst; q++) std::copy(t->begin(), t->begin() + bs, t->begin() + q * bs);
  }
  h->mat_dirty = true;
}

// omega.w of a particle: its type plus its instance above kTypeBits
int type_bits(const int *type, const std::vector<int> &inst_of, int64_t i) {
  return type[i] + (inst_of.empty() ? 0 : (inst_of[i] << dem::kTypeBits));
}

// history values per pair / plane-wall / mesh contact: the shear (3), and the
// rolling torque beside it (3 more) when some model is epsd2. Models, walls
// and meshes are fixed before the particles, which allocate the histories.
int needed_hist_width(const dem_handle *h) {
  bool epsd2 = (model_bits(h->pair_model) & dem::kModelEpsd2) != 0;
  for (const Wall &w : h->walls) epsd2 = epsd2 || (w.model & dem::kModelEpsd2);
  for (const auto &m : h->meshes) epsd2 = epsd2 || (m.model & dem::kModelEpsd2);
  return epsd2 ? 6 : 3;
}

// Uploads the material tables the step kernel reads and checks that every
// model finds its properties. need_particles = false: the set-up of a system
// whose particles are still to be inserted.
int check_ready(dem_handle *h, bool need_particles = true) {
  if (!h) return DEM_ERR_ARG;
  if (!h->mat_set) return fail(h, DEM_ERR_STATE, "materials not set (fix property/global)");
  if (need_particles && h->n <= 0) return fail(h, DEM_ERR_STATE, "no particles");
  int rc = upload_materials(h);
  if (rc) return rc;
  for (const Wall &w : h->walls)
    if (w.type < 1 || w.type > h->ntypes)
      return fail(h, DEM_ERR_ARG, "wall type out of range of the material tables");
  // every model in use finds its properties (PropertyRegistry::connect
  // fails the same way when a fix property/global is missing)
  std::vector<int> models{model_bits(h->pair_model)};
  for (const Wall &w : h->walls) models.push_back(w.model);
  for (const auto &m : h->meshes) models.push_back(m.model);
  for (int b : models) {
    const int nm = (b >> dem::kModelNormalShift) & 3;
    const bool elastic = nm == dem::kNormalHertz || nm == dem::kNormalHooke;
    if (elastic && !h->have_elastic)
      return fail(h, DEM_ERR_ARG,
                  std::string("model ") + normal_model_name(nm) +
                      " needs youngsModulus, poissonsRatio and coefficientRestitution");
    if (!elastic && !h->have_stiff)
      return fail(h, DEM_ERR_ARG,
                  std::string("model ") + normal_model_name(nm) + " needs kn, kt, gamman and gammat");
    if (nm == dem::kNormalHooke && h->char_vel < 1e-2)  // global_properties.cpp:233-237
      return fail(h, DEM_ERR_ARG, "characteristicVelocity >= 1e-2 required for SI units");
    if ((b & dem::kModelSjkr) && !h->have_ced)
      return fail(h, DEM_ERR_ARG, "cohesion sjkr needs cohesionEnergyDensity");
    // the rolling torque is kept beside the tangential history
    // (rolling_model_epsd2.h:172)
    if ((b & dem::kModelEpsd2) && (b & dem::kModelNoHistory))
      return fail(h, DEM_ERR_ARG, "rolling_friction epsd2 needs tangential history");
    if ((b & dem::kModelEpsd2) && (b & 1))
      return fail(h, DEM_ERR_ARG, "rolling_friction: cdt and epsd2 are exclusive");
  }
  if (h->n > 0 && needed_hist_width(h) != h->hw)
    return fail(h, DEM_ERR_STATE, "contact history width out of date");
  return DEM_OK;
}

int do_setup(dem_handle *h) {
  if (h && h->n == 0) {
    // Verlet::setup of a system whose particles are still to be inserted:
    // nothing to build or compute yet, but the material tables go to the
    // device now (the particles that join later are stepped with them)
    int rc = check_ready(h, false);
    if (rc) return rc;
    h->setup_done = true;
    h->can_continue = true;
    return DEM_OK;
  }
  int rc = check_ready(h);
  if (rc) return rc;
  HIPCHK(h, hipSetDevice(h->p.device));
  h->zero_vnode = true;  // FixMoveMesh::setup zeroes node velocities (fix_move_mesh.cpp:211-216)
  rc = rebuild(h);
  h->zero_vnode = false;
  if (rc) return rc;
  if ((rc = force_pass(h, 0))) return rc;  // pair_gran.cpp:665-667: shearupdate 0 in setup
  h->setup_done = true;
  h->can_continue = true;
  return DEM_OK;
}

// Re-allocates every per-particle array with room for newcap particles. The
// arrays are column-major with the capacity as column stride; the first n
// entries of every column are kept, new memory is zero (mesh slots: empty).
int grow_particles(dem_handle *h, int newcap) {
  const int oc = h->cap;
  const int64_t n = h->n;
  const int nw = (int)std::max<size_t>(1, h->walls.size());
  const int nm = (int)std::max<size_t>(1, h->meshes.size());
  const size_t osz = h->mixed ? sizeof(float4) : sizeof(double4);
  struct Col {
    void **p;
    size_t elem;
    int cols;
    int fill;   // byte value of new memory
    bool keep;  // copy the first n entries of every column
  };
  std::vector<Col> cols;
  for (int b = 0; b < 2; b++) {
    cols.push_back({(void **)&h->st[b].pos, sizeof(double4), 1, 0, true});
    cols.push_back({(void **)&h->st[b].vel, sizeof(double4), 1, 0, true});
    cols.push_back({&h->st[b].omg, osz, 1, 0, true});
    cols.push_back({(void **)&h->aux[b].tag, sizeof(int), 1, 0, true});
    cols.push_back({(void **)&h->aux[b].orig, sizeof(int), 1, 0, true});
    cols.push_back({(void **)&h->aux[b].meta, sizeof(int), 1, 0, true});
    cols.push_back({&h->aux[b].whist, h->fsz, h->hw * nw, 0, true});
    cols.push_back({(void **)&h->aux[b].mslot, sizeof(int), nm * dem::kMeshSlots, 0xff, true});
    cols.push_back({&h->aux[b].mhist, h->fsz, h->hw * nm * dem::kMeshSlots, 0, true});
    cols.push_back({(void **)&h->lst[b].nbr, sizeof(uint32_t), h->W, 0, true});
    cols.push_back({&h->lst[b].hist, h->fsz * dem::hist_rec(h->hw), h->W, 0, true});
  }
  cols.push_back({(void **)&h->xhold, sizeof(double), 3, 0, true});
  cols.push_back({(void **)&h->fout, sizeof(double), 6, 0, true});
  cols.push_back({(void **)&h->key, sizeof(int), 1, 0, false});
  cols.push_back({(void **)&h->perm, sizeof(int), 1, 0, false});
  cols.push_back({(void **)&h->d_cand, sizeof(int), dem::kMaxMeshCand, 0, true});
  cols.push_back({(void **)&h->d_ncand, sizeof(int), 1, 0, true});
  int rc;
  for (Col &c : cols) {
    void *np = nullptr;
    const size_t bytes = c.elem * (size_t)c.cols * (size_t)newcap;
    if ((rc = dalloc(h, &np, bytes))) return rc;
    HIPCHK(h, hipMemsetAsync(np, c.fill, bytes, h->stream));
    if (c.keep && n > 0 && *c.p)
      HIPCHK(h, hipMemcpy2DAsync(np, c.elem * newcap, *c.p, c.elem * oc, c.elem * n, c.cols,
                                 hipMemcpyDeviceToDevice, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    dfree(h, *c.p);
    *c.p = np;
  }
  h->cap = newcap;
  clear_graphs(h);
  return DEM_OK;
}

// Particles of dem_insert_particles join the system: appended behind the
// present ones in the current buffers, with no list, wall or mesh history.
// The caller forces a neighbor build before their first force pass.
int append_pending(dem_handle *h) {
  dem_handle::Pending q;
  std::swap(q, h->pend);
  const int64_t m = (int64_t)q.tag.size();
  if (m == 0) return DEM_OK;
  if (h->n == 0) {  // the handle's first particles
    int rc = dem_set_particles(h, m, q.x.data(), q.v.data(), q.w.data(), q.r.data(), q.m.data(),
                               q.type.data(), q.tag.data());
    if (rc) return rc;
    // the step kernel reads the material tables: make sure they are on the
    // device whatever happened since the set-up
    return check_ready(h);
  }
  const int64_t n0 = h->n, n1 = n0 + m;
  if (n1 > 0x7fffffffLL / 2) return fail(h, DEM_ERR_ARG, "too many particles");
  int rc;
  if (n1 > h->cap && (rc = grow_particles(h, (int)std::max<int64_t>(n1, (int64_t)h->cap + h->cap / 2))))
    return rc;
  std::vector<double4> pos(m), vel(m);
  std::vector<int> orig(m), zero(m, 0);
  for (int64_t k = 0; k < m; k++) {
    pos[k] = make_double4(q.x[3 * k], q.x[3 * k + 1], q.x[3 * k + 2], q.r[k]);
    vel[k] = make_double4(q.v[3 * k], q.v[3 * k + 1], q.v[3 * k + 2], q.m[k]);
    orig[k] = (int)(n0 + k);
  }
  const StateBuf &st = h->st[h->cur_st];
  HIPCHK(h, hipMemcpyAsync(st.pos + n0, pos.data(), sizeof(double4) * m, hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipMemcpyAsync(st.vel + n0, vel.data(), sizeof(double4) * m, hipMemcpyHostToDevice, h->stream));
  std::vector<float4> wf;
  std::vector<double4> wd;
  if (h->mixed) {
    wf.resize(m);
    for (int64_t k = 0; k < m; k++)
      wf[k] = make_float4((float)q.w[3 * k], (float)q.w[3 * k + 1], (float)q.w[3 * k + 2], (float)q.type[k]);
    HIPCHK(h, hipMemcpyAsync((float4 *)st.omg + n0, wf.data(), sizeof(float4) * m, hipMemcpyHostToDevice,
                             h->stream));
  } else {
    wd.resize(m);
    for (int64_t k = 0; k < m; k++)
      wd[k] = make_double4(q.w[3 * k], q.w[3 * k + 1], q.w[3 * k + 2], (double)q.type[k]);
    HIPCHK(h, hipMemcpyAsync((double4 *)st.omg + n0, wd.data(), sizeof(double4) * m, hipMemcpyHostToDevice,
                             h->stream));
  }
  const AuxBuf &ax = h->aux[h->cur_aux];
  HIPCHK(h, hipMemcpyAsync(ax.tag + n0, q.tag.data(), sizeof(int) * m, hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipMemcpyAsync(ax.orig + n0, orig.data(), sizeof(int) * m, hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipMemcpyAsync(ax.meta + n0, zero.data(), sizeof(int) * m, hipMemcpyHostToDevice, h->stream));
  const int nm = (int)std::max<size_t>(1, h->meshes.size());
  for (int b = 0; b < 2; b++)
    HIPCHK(h, hipMemset2DAsync(h->aux[b].mslot + n0, sizeof(int) * (size_t)h->cap, 0xff, sizeof(int) * m,
                               (size_t)nm * dem::kMeshSlots, h->stream));
  HIPCHK(h, hipMemsetAsync(h->d_ncand + n0, 0, sizeof(int) * m, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  h->n = n1;
  clear_graphs(h);
  // a larger particle may need larger cells
  for (int64_t k = 0; k < m; k++) h->rmax = std::max(h->rmax, q.r[k]);
  const double rm = std::max(h->rmax, h->rmax_future);
  if ((rm + rm) * h->p.cdf + h->p.skin != h->cutneighmax) {
    for (int **p : {&h->cell_rank, &h->cell_count, &h->cell_start, &h->scan_tmp}) {
      dfree(h, *p);
      *p = nullptr;
    }
    if ((rc = setup_grid(h))) return rc;
    if (h->ntri > 0) {
      dfree(h, h->d_tcell_count);
      dfree(h, h->d_tcell_start);
      if ((rc = dalloc(h, (void **)&h->d_tcell_count, sizeof(int) * h->ncell))) return rc;
      if ((rc = dalloc(h, (void **)&h->d_tcell_start, sizeof(int) * (h->ncell + 1)))) return rc;
    }
  }
  return DEM_OK;
}

}  // namespace

extern "C" {

const char *dem_last_error(const dem_handle *h) {
  return h ? h->err.c_str() : g_create_error.c_str();
}

int dem_create(const dem_params *params, dem_handle **out) {
  if (!params || !out) return fail(nullptr, DEM_ERR_ARG, "null argument");
  *out = nullptr;
  const dem_params &p = *params;
  if (!(p.dt > 0.0)) return fail(nullptr, DEM_ERR_ARG, "timestep must be > 0");
  if (!(p.skin >= 0.0)) return fail(nullptr, DEM_ERR_ARG, "skin must be >= 0");
  if (!(p.cdf >= 1.0)) return fail(nullptr, DEM_ERR_ARG, "contact_distance_factor must be >= 1");
  for (int d = 0; d < 3; d++)
    if (!(p.boxhi[d] > p.boxlo[d])) return fail(nullptr, DEM_ERR_ARG, "empty box");
  if (p.precision != DEM_FP64 && p.precision != DEM_MIXED)
    return fail(nullptr, DEM_ERR_ARG, "unknown precision mode");
  if (p.neigh_every < 0) return fail(nullptr, DEM_ERR_ARG, "neigh_modify every must be > 0");
  for (int d = 0; d < 3; d++)
    if (p.periodic[d] && ((p.shrink_faces >> (2 * d)) & 3))
      return fail(nullptr, DEM_ERR_ARG, "Both sides of boundary must be periodic");
  if ((p.shrink_min & ~p.shrink_faces) || (p.shrink_faces & ~0x3f))
    return fail(nullptr, DEM_ERR_ARG, "bad shrink-wrapped face bits");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= p.device || p.device < 0)
    return fail(nullptr, DEM_ERR_NODEV, "no HIP device available");
  if (hipSetDevice(p.device) != hipSuccess) return fail(nullptr, DEM_ERR_NODEV, "hipSetDevice failed");
  dem_handle *h = new dem_handle();
  h->p = p;
  h->pair_model.rolling_cdt = p.rolling_cdt;
  h->pair_model.tangential_damping = p.tangential_damping;
  h->pair_model.limit_force = p.limit_force;
  h->pair_model.torsion_torque = p.torsion_torque;
  {
    int ncu = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, p.device) == hipSuccess &&
        ncu > 0)
      h->ncu = ncu;
  }
  h->mixed = p.precision == DEM_MIXED;
  h->fsz = h->mixed ? 4 : 8;
  h->W = p.max_neighbors > 0 ? p.max_neighbors : 32;
  if (h->W > 0xffff) {
    delete h;
    return fail(nullptr, DEM_ERR_ARG, "max_neighbors too large");
  }
  if (hipStreamCreate(&h->stream) != hipSuccess) {
    delete h;
    return fail(nullptr, DEM_ERR_HIP, "hipStreamCreate failed");
  }
  if (dalloc(h, (void **)&h->ctrl, sizeof(Ctrl)) ||
      hipHostMalloc((void **)&h->ctrl_host, sizeof(Ctrl), hipHostMallocDefault) != hipSuccess) {
    g_create_error = h->err.empty() ? "pinned allocation failed" : h->err;
    dem_destroy(h);
    return DEM_ERR_HIP;
  }
  if (hipHostMalloc((void **)&h->h_total, sizeof(int), hipHostMallocDefault) != hipSuccess) {
    g_create_error = "pinned allocation failed";
    dem_destroy(h);
    return DEM_ERR_HIP;
  }
  std::memset(h->ctrl_host, 0, sizeof(Ctrl));
  h->ctrl_host->stop_at = kNever;
  hipMemcpy(h->ctrl, h->ctrl_host, sizeof(Ctrl), hipMemcpyHostToDevice);
  h->ev.resize(2);
  for (auto &e : h->ev) hipEventCreate(&e);
  *out = h;
  return DEM_OK;
}

int dem_destroy(dem_handle *h) {
  if (!h) return DEM_OK;
  if (h->stream) hipStreamSynchronize(h->stream);
  clear_graphs(h);
  for (void *p : h->allocs) hipFree(p);
  for (auto &e : h->ev)
    if (e) hipEventDestroy(e);
  if (h->ctrl_host) hipHostFree(h->ctrl_host);
  if (h->h_total) hipHostFree(h->h_total);
  if (h->stream) hipStreamDestroy(h->stream);
  delete h;
  return DEM_OK;
}

int dem_set_materials(dem_handle *h, int ntypes, const double *Y, const double *nu,
                      const double *e, const double *mu, const double *mur) {
  if (!h) return DEM_ERR_ARG;
  int rc = check_material_args(h, ntypes, Y, nu, e, mu);
  if (rc) return rc;
  if (h->n > 0 && ntypes < h->ntypes)
    return fail(h, DEM_ERR_STATE, "fewer atom types than the particles use");
  h->ntypes = ntypes;
  const size_t bs = (size_t)(ntypes + 1) * (ntypes + 1);
  for (auto *t : host_tables(h)) t->assign(bs, 0.0);
  mix_block(h, 0, Y, nu, e, mu, mur);
  replicate_material_blocks(h);  // every instance starts with these materials
  h->have_elastic = Y != nullptr;
  h->have_stiff = h->have_ced = false;  // dem_set_model_coefficients comes after
  h->mat_set = true;
  h->setup_done = false;
  h->can_continue = false;
  return DEM_OK;
}

int dem_set_pair_model(dem_handle *h, const dem_contact_model *model) {
  if (!h || !model) return DEM_ERR_ARG;
  if (model->normal_model < 0 || model->normal_model > 3 || model->tangential_model < 0 ||
      model->tangential_model > 1 || model->cohesion_model < 0 || model->cohesion_model > 1)
    return fail(h, DEM_ERR_ARG, "unknown contact model");
  if (h->n > 0) return fail(h, DEM_ERR_STATE, "set the pair model before particles");
  h->pair_model = *model;
  h->setup_done = false;
  h->can_continue = false;
  return DEM_OK;
}

int dem_set_model_coefficients(dem_handle *h, int ntypes, const double *kn, const double *kt,
                               const double *gamman, const double *gammat,
                               const double *cohesion_energy_density, double characteristic_velocity) {
  if (!h) return DEM_ERR_ARG;
  if (!h->mat_set) return fail(h, DEM_ERR_STATE, "call dem_set_materials first");
  if (ntypes != h->ntypes) return fail(h, DEM_ERR_ARG, "ntypes differs from the materials'");
  if ((kn || kt || gamman || gammat) && !(kn && kt && gamman && gammat))
    return fail(h, DEM_ERR_ARG, "kn, kt, gamman and gammat go together");
  const int t1 = ntypes + 1;
  const double *src[5] = {kn, kt, gamman, gammat, cohesion_energy_density};
  std::vector<double> *dst[5] = {&h->hkn, &h->hkt, &h->hgn, &h->hgt, &h->hced};
  for (int q = 0; q < 5; q++) {
    if (!src[q]) continue;
    for (int i = 1; i <= ntypes; i++)
      for (int j = 1; j <= ntypes; j++) (*dst[q])[i * t1 + j] = src[q][(i - 1) * ntypes + (j - 1)];
  }
  if (kn) h->have_stiff = true;
  if (cohesion_energy_density) h->have_ced = true;
  h->char_vel = characteristic_velocity;
  replicate_material_blocks(h);
  h->setup_done = false;
  h->can_continue = false;
  return DEM_OK;
}

int dem_set_instances(dem_handle *h, int ninst, const int64_t *counts) {
  if (!h) return DEM_ERR_ARG;
  if (ninst < 1 || ninst > dem::kMaxInstances || !counts)
    return fail(h, DEM_ERR_ARG, "bad instance arguments");
  if (h->n > 0) return fail(h, DEM_ERR_STATE, "set instances before particles");
  for (int q = 0; q < ninst; q++)
    if (counts[q] <= 0) return fail(h, DEM_ERR_ARG, "every instance needs particles");
  h->ninst = ninst;
  h->inst_count.assign(counts, counts + ninst);
  if (h->mat_set) replicate_material_blocks(h);
  return DEM_OK;
}

int dem_set_instance_materials(dem_handle *h, int inst, int ntypes, const double *Y,
                               const double *nu, const double *e, const double *mu,
                               const double *mur) {
  if (!h) return DEM_ERR_ARG;
  if (!h->mat_set) return fail(h, DEM_ERR_STATE, "call dem_set_materials first");
  if (inst < 0 || inst >= h->ninst) return fail(h, DEM_ERR_ARG, "bad instance index");
  if (ntypes != h->ntypes) return fail(h, DEM_ERR_ARG, "every instance has the same atom types");
  int rc = check_material_args(h, ntypes, Y, nu, e, mu);
  if (rc) return rc;
  mix_block(h, inst, Y, nu, e, mu, mur);
  h->mat_dirty = true;
  h->setup_done = false;
  h->can_continue = false;
  return DEM_OK;
}

int dem_get_instance_stats(dem_handle *h, int ninst, int64_t *ncontacts, int64_t *nwall) {
  if (!h || ninst < 1) return DEM_ERR_ARG;
  if (ninst > h->ninst) return fail(h, DEM_ERR_ARG, "more instances asked for than the system has");
  std::vector<unsigned long long> st(2 * (size_t)h->ninst, 0ull);
  if (h->d_inst_stats) {
    HIPCHK(h, hipStreamSynchronize(h->stream));
    HIPCHK(h, hipMemcpy(st.data(), h->d_inst_stats, sizeof(unsigned long long) * st.size(),
                        hipMemcpyDeviceToHost));
  } else if (h->ninst == 1) {
    st[0] = (unsigned long long)h->ncontacts * 2;
    st[1] = (unsigned long long)h->nwall_contacts;
  }
  for (int q = 0; q < ninst; q++) {
    if (ncontacts) ncontacts[q] = (int64_t)(st[q] / 2);
    if (nwall) nwall[q] = (int64_t)st[h->ninst + q];
  }
  return DEM_OK;
}

int dem_add_plane_wall(dem_handle *h, int axis, double position, int wall_type,
                       const dem_contact_model *model) {
  if (!h) return DEM_ERR_ARG;
  if (axis < 0 || axis > 2) return fail(h, DEM_ERR_ARG, "wall axis must be 0, 1 or 2");
  if ((int)h->walls.size() >= dem::kMaxWalls) return fail(h, DEM_ERR_ARG, "too many walls");
  if (h->n > 0) return fail(h, DEM_ERR_STATE, "add walls before particles");
  dem_contact_model m = h->pair_model;
  if (model) m = *model;
  h->walls.push_back({axis, position, wall_type, model_bits(m)});
  return DEM_OK;
}

int dem_add_mesh(dem_handle *h, int64_t ntri, const double *nodes, int wall_type,
                 const dem_contact_model *model, int *mesh_id) {
  if (!h) return DEM_ERR_ARG;
  if (ntri <= 0 || ntri > (1 << 24) || !nodes) return fail(h, DEM_ERR_ARG, "bad mesh arguments");
  if ((int)h->meshes.size() >= dem::kMaxMeshes) return fail(h, DEM_ERR_ARG, "too many meshes");
  if (h->n > 0) return fail(h, DEM_ERR_STATE, "add meshes before particles");
  dem::HostMesh hm;
  std::string err;
  if (!dem::build_host_mesh(hm, (int)ntri, nodes, h->p.boxlo, h->p.boxhi, err))
    return fail(h, DEM_ERR_ARG, err);
  hm.wall_type = wall_type;
  dem_contact_model m = h->pair_model;
  if (model) m = *model;
  hm.model = model_bits(m);
  h->meshes.push_back(std::move(hm));
  h->ntri += h->meshes.back().ntri;
  if (mesh_id) *mesh_id = (int)h->meshes.size() - 1;
  return DEM_OK;
}

int dem_mesh_generate(const char *type, const double *dims, int nsegments, int mask, int64_t cap, double *nodes,
                      int64_t *ntri) {
  if (!type || !dims || !ntri) return fail(nullptr, DEM_ERR_ARG, "null argument");
  *ntri = 0;
  const std::string ty(type);
  const int nd = ty == "box" ? 3 : ty == "pipe" ? 3 : ty == "cylinder" ? 2 : 1;
  for (int k = 0; k < nd; k++)
    if (!(dims[k] > 0)) return fail(nullptr, DEM_ERR_ARG, "generator sizes must be > 0");
  const bool segs = ty == "cylinder" || ty == "pipe" || ty == "disk";
  const bool masked = ty == "cube" || ty == "box" || ty == "cylinder" || ty == "pipe";
  if ((segs && nsegments <= 0) || (masked && mask <= 0))
    return fail(nullptr, DEM_ERR_ARG, "nsegments and mask must be > 0");
  double dv[3] = {dims[0], nd > 1 ? dims[1] : 0.0, nd > 2 ? dims[2] : 0.0};
  std::vector<double> out;
  std::string err;
  if (!dem::generate_mesh(ty, dv, nsegments, mask, out, err)) return fail(nullptr, DEM_ERR_ARG, err);
  *ntri = (int64_t)(out.size() / 9);
  if (nodes) std::copy(out.begin(), out.begin() + 9 * std::min<int64_t>(cap, *ntri), nodes);
  return DEM_OK;
}

int dem_mesh_rotate(dem_handle *h, int mesh_id, const double *origin, const double *axis,
                    double period) {
  if (!h) return DEM_ERR_ARG;
  if (mesh_id < 0 || mesh_id >= (int)h->meshes.size() || !origin || !axis || !(period != 0.0))
    return fail(h, DEM_ERR_ARG, "bad move/mesh rotate arguments");
  if (h->n > 0) return fail(h, DEM_ERR_STATE, "set mesh motion before particles");
  dem::HostMesh &m = h->meshes[mesh_id];
  if (m.moving && m.mover != dem::kMoverRotate)
    return fail(h, DEM_ERR_ARG, "one fix move/mesh per mesh on this path");
  // MeshMoverRotate: axis normalized, omega = 2 pi / period (mesh_mover_rotation.cpp:66-93)
  const double norm = std::sqrt(axis[0] * axis[0] + axis[1] * axis[1] + axis[2] * axis[2]);
  if (norm == 0.) return fail(h, DEM_ERR_ARG, "rotation axis is zero");
  const double invnorm = 1. / norm;
  for (int k = 0; k < 3; k++) {
    m.origin[k] = origin[k];
    m.axis[k] = axis[k] * invnorm;
  }
  m.omega = 2. * kMyPi / period;
  m.moving = true;
  m.mover = dem::kMoverRotate;
  return DEM_OK;
}

int dem_mesh_stress(dem_handle *h, int mesh_id, const double *reference_point) {
  if (!h) return DEM_ERR_ARG;
  if (mesh_id < 0 || mesh_id >= (int)h->meshes.size()) return fail(h, DEM_ERR_ARG, "bad mesh id");
  dem::HostMesh &m = h->meshes[mesh_id];
  m.stress = true;
  for (int k = 0; k < 3; k++) m.p_ref[k] = reference_point ? reference_point[k] : 0.0;
  clear_graphs(h);  // the flag travels in the captured kernel arguments
  return DEM_OK;
}

int dem_get_mesh_force(dem_handle *h, int mesh_id, double *force, double *torque) {
  if (!h) return DEM_ERR_ARG;
  if (mesh_id < 0 || mesh_id >= (int)h->meshes.size()) return fail(h, DEM_ERR_ARG, "bad mesh id");
  if (!h->meshes[mesh_id].stress) return fail(h, DEM_ERR_STATE, "the mesh does not track its stress");
  int rc = read_ctrl(h);
  if (rc) return rc;
  for (int k = 0; k < 3; k++) {
    if (force) force[k] = h->ctrl_host->mesh_force[mesh_id][k];
    if (torque) torque[k] = h->ctrl_host->mesh_force[mesh_id][3 + k];
  }
  return DEM_OK;
}

int dem_mesh_linear(dem_handle *h, int mesh_id, const double *velocity) {
  if (!h) return DEM_ERR_ARG;
  if (mesh_id < 0 || mesh_id >= (int)h->meshes.size() || !velocity)
    return fail(h, DEM_ERR_ARG, "bad move/mesh linear arguments");
  if (h->n > 0) return fail(h, DEM_ERR_STATE, "set mesh motion before particles");
  dem::HostMesh &m = h->meshes[mesh_id];
  if (m.moving) return fail(h, DEM_ERR_ARG, "one fix move/mesh per mesh on this path");
  for (int k = 0; k < 3; k++) m.vel[k] = velocity[k];  // MeshMoverLinear (mesh_mover_linear.cpp:62-75)
  m.moving = true;
  m.mover = dem::kMoverLinear;
  return DEM_OK;
}

int dem_mesh_viblin(dem_handle *h, int mesh_id, const double *axis, int order, const double *amplitude,
                    const double *phase, const double *period) {
  if (!h) return DEM_ERR_ARG;
  if (mesh_id < 0 || mesh_id >= (int)h->meshes.size() || !axis || !amplitude || !phase || !period)
    return fail(h, DEM_ERR_ARG, "bad move/mesh viblin arguments");
  if (order < 1 || order > 30) return fail(h, DEM_ERR_ARG, "order can be at most 30 and must be greater 0");
  if (h->n > 0) return fail(h, DEM_ERR_STATE, "set mesh motion before particles");
  dem::HostMesh &m = h->meshes[mesh_id];
  if (m.moving) return fail(h, DEM_ERR_ARG, "one fix move/mesh per mesh on this path");
  // MeshMoverVibLin (mesh_mover_linear.cpp:230-263): axis normalized, omega_j = 2 pi / period_j
  const double norm = std::sqrt(axis[0] * axis[0] + axis[1] * axis[1] + axis[2] * axis[2]);
  const double invnorm = (norm == 0.) ? 0. : 1. / norm;
  for (int k = 0; k < 3; k++) m.axis[k] = axis[k] * invnorm;
  m.vord = order;
  for (int j = 0; j < order; j++) {
    m.vampl[j] = amplitude[j];
    m.vphi[j] = phase[j];
    m.vomega[j] = 2. * kMyPi / period[j];
  }
  m.moving = true;
  m.mover = dem::kMoverVibLin;
  return DEM_OK;
}

int dem_mesh_vibrot(dem_handle *h, int mesh_id, const double *origin, const double *axis, int order,
                    const double *amplitude, const double *phase, const double *period) {
  if (!h) return DEM_ERR_ARG;
  if (mesh_id < 0 || mesh_id >= (int)h->meshes.size() || !origin || !axis || !amplitude || !phase ||
      !period)
    return fail(h, DEM_ERR_ARG, "bad move/mesh vibrot arguments");
  if (order < 1 || order > 30) return fail(h, DEM_ERR_ARG, "order can be at most 30 and must be greater 0");
  if (h->n > 0) return fail(h, DEM_ERR_STATE, "set mesh motion before particles");
  dem::HostMesh &m = h->meshes[mesh_id];
  if (m.moving) return fail(h, DEM_ERR_ARG, "one fix move/mesh per mesh on this path");
  // MeshMoverVibRot (mesh_mover_rotation.cpp:333-375)
  const double norm = std::sqrt(axis[0] * axis[0] + axis[1] * axis[1] + axis[2] * axis[2]);
  const double invnorm = (norm == 0.) ? 0. : 1. / norm;
  for (int k = 0; k < 3; k++) {
    m.origin[k] = origin[k];
    m.axis[k] = axis[k] * invnorm;
  }
  m.vord = order;
  for (int j = 0; j < order; j++) {
    m.vampl[j] = amplitude[j];
    m.vphi[j] = phase[j];
    m.vomega[j] = 2. * kMyPi / period[j];
  }
  m.moving = true;
  m.mover = dem::kMoverVibRot;
  return DEM_OK;
}

int dem_mesh_wiggle(dem_handle *h, int mesh_id, const double *amplitude, double period) {
  if (!h) return DEM_ERR_ARG;
  if (mesh_id < 0 || mesh_id >= (int)h->meshes.size() || !amplitude || !(period != 0.0))
    return fail(h, DEM_ERR_ARG, "bad move/mesh wiggle arguments");
  if (h->n > 0) return fail(h, DEM_ERR_STATE, "set mesh motion before particles");
  dem::HostMesh &m = h->meshes[mesh_id];
  if (m.moving) return fail(h, DEM_ERR_ARG, "one fix move/mesh per mesh on this path");
  // MeshMoverWiggle (mesh_mover_linear.cpp:160-183): omega = 2 pi / period
  for (int k = 0; k < 3; k++) m.ampl[k] = amplitude[k];
  m.womega = 2. * kMyPi / period;
  m.moving = true;
  m.mover = dem::kMoverWiggle;
  return DEM_OK;
}

int dem_get_mesh_nodes(dem_handle *h, int mesh_id, double *nodes, double *vnodes) {
  if (!h) return DEM_ERR_ARG;
  if (mesh_id < 0 || mesh_id >= (int)h->meshes.size()) return fail(h, DEM_ERR_ARG, "bad mesh id");
  int t0 = 0;
  for (int m = 0; m < mesh_id; m++) t0 += h->meshes[m].ntri;
  const int T = h->meshes[mesh_id].ntri;
  if (!h->d_tri_buf[0]) {  // not uploaded yet: host copy
    for (int t = 0; t < T; t++)
      for (int k = 0; k < 9; k++) {
        if (nodes) nodes[9 * t + k] = h->meshes[mesh_id].node[9 * t + k];
        if (vnodes) vnodes[9 * t + k] = 0.0;
      }
    return DEM_OK;
  }
  HIPCHK(h, hipStreamSynchronize(h->stream));
  std::vector<dem::TriGeom> g(T);
  HIPCHK(h, hipMemcpy(g.data(), h->d_tri_buf[h->tri_cur] + t0, sizeof(dem::TriGeom) * T,
                      hipMemcpyDeviceToHost));
  for (int t = 0; t < T; t++)
    for (int k = 0; k < 9; k++) {
      if (nodes) nodes[9 * t + k] = g[t].n[k];
      if (vnodes) vnodes[9 * t + k] = g[t].vn[k];
    }
  return DEM_OK;
}

int dem_get_mesh_topology(dem_handle *h, int mesh_id, int *edge_active, int *corner_active,
                          int *obtuse) {
  if (!h) return DEM_ERR_ARG;
  if (mesh_id < 0 || mesh_id >= (int)h->meshes.size()) return fail(h, DEM_ERR_ARG, "bad mesh id");
  const dem::HostMesh &m = h->meshes[mesh_id];
  for (int t = 0; t < m.ntri; t++) {
    for (int k = 0; k < 3; k++) {
      if (edge_active) edge_active[3 * t + k] = m.eact[3 * t + k];
      if (corner_active) corner_active[3 * t + k] = m.cact[3 * t + k];
    }
    if (obtuse) obtuse[t] = m.obtuse[t];
  }
  return DEM_OK;
}

int dem_get_mesh_history(dem_handle *h, int mesh_id, int maxslot, int *tri, double *hist) {
  if (!h || !tri || !hist || maxslot <= 0) return DEM_ERR_ARG;
  if (mesh_id < 0 || mesh_id >= (int)h->meshes.size()) return fail(h, DEM_ERR_ARG, "bad mesh id");
  if (h->n <= 0) return fail(h, DEM_ERR_STATE, "no particles");
  HIPCHK(h, hipStreamSynchronize(h->stream));
  const int64_t n = h->n;
  const size_t cap = (size_t)h->cap;
  int t0 = 0;
  for (int m = 0; m < mesh_id; m++) t0 += h->meshes[m].ntri;
  const dem::AuxBuf &a = h->aux[h->cur_aux];
  std::vector<int> slot((size_t)dem::kMeshSlots * cap), orig(n);
  std::vector<char> raw(h->fsz * h->hw * dem::kMeshSlots * cap);
  HIPCHK(h, hipMemcpy(slot.data(), a.mslot + (size_t)mesh_id * dem::kMeshSlots * cap,
                      sizeof(int) * slot.size(), hipMemcpyDeviceToHost));
  HIPCHK(h, hipMemcpy(raw.data(), (char *)a.mhist + h->fsz * h->hw * dem::kMeshSlots * cap * mesh_id,
                      raw.size(), hipMemcpyDeviceToHost));
  HIPCHK(h, hipMemcpy(orig.data(), a.orig, sizeof(int) * n, hipMemcpyDeviceToHost));
  for (int64_t i = 0; i < n; i++) {
    const int o = orig[i];
    for (int k = 0; k < maxslot; k++) {
      const bool inside = k < dem::kMeshSlots;
      const int st = inside ? slot[(size_t)k * cap + i] : -1;
      tri[(size_t)o * maxslot + k] = st >= 0 ? st - t0 : -1;
      for (int d = 0; d < 3; d++) {
        double val = 0.0;
        if (st >= 0) {
          const size_t idx = (size_t)(k * h->hw + d) * cap + i;
          val = h->mixed ? (double)((float *)raw.data())[idx] : ((double *)raw.data())[idx];
        }
        hist[((size_t)o * maxslot + k) * 3 + d] = val;
      }
    }
  }
  return DEM_OK;
}

int dem_set_particles(dem_handle *h, int64_t n, const double *x, const double *v,
                      const double *omega, const double *radius, const double *rmass,
                      const int *type, const int *tag) {
  if (!h) return DEM_ERR_ARG;
  if (n <= 0 || n > 0x7fffffffLL / 2 || !x || !v || !omega || !radius || !rmass || !type || !tag)
    return fail(h, DEM_ERR_ARG, "bad particle arrays");
  if (!h->mat_set) return fail(h, DEM_ERR_STATE, "set materials before particles");
  HIPCHK(h, hipSetDevice(h->p.device));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  double rmax = 0.0;
  for (int64_t i = 0; i < n; i++) {
    if (!(radius[i] > 0.0)) return fail(h, DEM_ERR_ARG, "Fix nve/sphere requires extended particles");
    if (!(rmass[i] > 0.0)) return fail(h, DEM_ERR_ARG, "particle mass must be > 0");
    if (type[i] < 1 || type[i] > h->ntypes) return fail(h, DEM_ERR_ARG, "Invalid atom type");
    if (tag[i] <= 0) return fail(h, DEM_ERR_ARG, "Invalid atom ID");
    for (int d = 0; d < 3; d++)
      if (!std::isfinite(x[3 * i + d]) || !std::isfinite(v[3 * i + d]) || !std::isfinite(omega[3 * i + d]))
        return fail(h, DEM_ERR_ARG, "non-finite particle state");
    rmax = std::max(rmax, radius[i]);
  }
  std::vector<int> inst_of;  // instance of each particle, caller order
  if (h->ninst > 1) {
    int64_t sum = 0;
    for (int64_t c : h->inst_count) sum += c;
    if (sum != n) return fail(h, DEM_ERR_ARG, "instance particle counts do not add up to n");
    inst_of.resize(n);
    int64_t i0 = 0;
    for (int q = 0; q < h->ninst; q++)
      for (int64_t k = 0; k < h->inst_count[q]; k++) inst_of[i0++] = q;
  }
  free_particles(h);
  h->hw = needed_hist_width(h);
  h->n = n;
  h->cap = (int)n;
  h->rmax = rmax;
  int rc;
  if ((rc = setup_grid(h))) return rc;
  const size_t cap = (size_t)h->cap;
  const size_t nw = std::max<size_t>(1, h->walls.size());
  for (int b = 0; b < 2; b++) {
    if ((rc = dalloc(h, (void **)&h->st[b].pos, sizeof(double4) * cap))) return rc;
    if ((rc = dalloc(h, (void **)&h->st[b].vel, sizeof(double4) * cap))) return rc;
    if ((rc = dalloc(h, &h->st[b].omg, (h->mixed ? sizeof(float4) : sizeof(double4)) * cap))) return rc;
    if ((rc = dalloc(h, (void **)&h->aux[b].tag, sizeof(int) * cap))) return rc;
    if ((rc = dalloc(h, (void **)&h->aux[b].orig, sizeof(int) * cap))) return rc;
    if ((rc = dalloc(h, (void **)&h->aux[b].meta, sizeof(int) * cap))) return rc;
    if ((rc = dalloc(h, &h->aux[b].whist, h->fsz * h->hw * nw * cap))) return rc;
    if ((rc = dalloc(h, (void **)&h->lst[b].nbr, sizeof(uint32_t) * h->W * cap))) return rc;
    if ((rc = dalloc(h, &h->lst[b].hist, h->fsz * dem::hist_rec(h->hw) * h->W * cap))) return rc;
  }
  if ((rc = dalloc(h, (void **)&h->xhold, sizeof(double) * 3 * cap))) return rc;
  if ((rc = dalloc(h, (void **)&h->fout, sizeof(double) * 6 * cap))) return rc;
  if ((rc = dalloc(h, (void **)&h->key, sizeof(int) * cap))) return rc;
  if ((rc = dalloc(h, (void **)&h->perm, sizeof(int) * cap))) return rc;
  const size_t nm = std::max<size_t>(1, h->meshes.size());
  for (int b = 0; b < 2; b++) {
    if ((rc = dalloc(h, (void **)&h->aux[b].mslot, sizeof(int) * nm * dem::kMeshSlots * cap)))
      return rc;
    if ((rc = dalloc(h, &h->aux[b].mhist, h->fsz * h->hw * nm * dem::kMeshSlots * cap))) return rc;
    HIPCHK(h, hipMemset(h->aux[b].mslot, 0xff, sizeof(int) * nm * dem::kMeshSlots * cap));
  }
  if (h->ntri > 0) {
    std::vector<dem::TriGeom> geo;
    std::vector<int> off(1, 0), idx;
    int t0 = 0;
    for (size_t m = 0; m < h->meshes.size(); m++) {
      dem::fill_tri_geom(h->meshes[m], (int)m, geo);
      for (int t = 0; t < h->meshes[m].ntri; t++) {
        for (int b : h->meshes[m].coplanar[t]) idx.push_back(b + t0);
        off.push_back((int)idx.size());
      }
      t0 += h->meshes[m].ntri;
    }
    if (idx.empty()) idx.push_back(-1);
    const int T = h->ntri;
    if ((rc = dalloc(h, (void **)&h->d_tri_buf[0], sizeof(dem::TriGeom) * T))) return rc;
    h->d_tri_buf[1] = h->d_tri_buf[0];
    if (any_moving_mesh(h) &&
        (rc = dalloc(h, (void **)&h->d_tri_buf[1], sizeof(dem::TriGeom) * T)))
      return rc;
    h->tri_cur = 0;
    if ((rc = dalloc(h, (void **)&h->d_node_last, sizeof(double) * 9 * T))) return rc;
    if ((rc = dalloc(h, (void **)&h->d_copl_off, sizeof(int) * (T + 1)))) return rc;
    if ((rc = dalloc(h, (void **)&h->d_copl_idx, sizeof(int) * idx.size()))) return rc;
    if ((rc = dalloc(h, (void **)&h->d_tcount, sizeof(int) * T))) return rc;
    if ((rc = dalloc(h, (void **)&h->d_toff, sizeof(int) * (T + 1)))) return rc;
    if ((rc = dalloc(h, (void **)&h->d_tcell_count, sizeof(int) * h->ncell))) return rc;
    if ((rc = dalloc(h, (void **)&h->d_tcell_start, sizeof(int) * (h->ncell + 1)))) return rc;
    if ((rc = dalloc(h, (void **)&h->d_total, sizeof(int)))) return rc;
    for (int b = 0; b < 2; b++)
      HIPCHK(h, hipMemcpy(h->d_tri_buf[b], geo.data(), sizeof(dem::TriGeom) * T,
                          hipMemcpyHostToDevice));
    HIPCHK(h, hipMemcpy(h->d_copl_off, off.data(), sizeof(int) * (T + 1), hipMemcpyHostToDevice));
    HIPCHK(h, hipMemcpy(h->d_copl_idx, idx.data(), sizeof(int) * idx.size(), hipMemcpyHostToDevice));
  }
  if ((rc = dalloc(h, (void **)&h->d_cand, sizeof(int) * dem::kMaxMeshCand * cap))) return rc;
  if ((rc = dalloc(h, (void **)&h->d_ncand, sizeof(int) * cap))) return rc;
  HIPCHK(h, hipMemset(h->d_ncand, 0, sizeof(int) * cap));
  h->d_inst_stats = nullptr;
  if (h->ninst > 1) {
    if ((rc = dalloc(h, (void **)&h->d_inst_stats, sizeof(unsigned long long) * 2 * h->ninst)))
      return rc;
    HIPCHK(h, hipMemset(h->d_inst_stats, 0, sizeof(unsigned long long) * 2 * h->ninst));
  }

  std::vector<double4> pos(n), vel(n);
  std::vector<int> orig(n), zero(n, 0);
  for (int64_t i = 0; i < n; i++) {
    pos[i] = make_double4(x[3 * i], x[3 * i + 1], x[3 * i + 2], radius[i]);
    vel[i] = make_double4(v[3 * i], v[3 * i + 1], v[3 * i + 2], rmass[i]);
    orig[i] = (int)i;
  }
  HIPCHK(h, hipMemcpy(h->st[0].pos, pos.data(), sizeof(double4) * n, hipMemcpyHostToDevice));
  HIPCHK(h, hipMemcpy(h->st[0].vel, vel.data(), sizeof(double4) * n, hipMemcpyHostToDevice));
  if (h->mixed) {
    std::vector<float4> w(n);
    for (int64_t i = 0; i < n; i++)
      w[i] = make_float4((float)omega[3 * i], (float)omega[3 * i + 1], (float)omega[3 * i + 2],
                         (float)type_bits(type, inst_of, i));
    HIPCHK(h, hipMemcpy(h->st[0].omg, w.data(), sizeof(float4) * n, hipMemcpyHostToDevice));
  } else {
    std::vector<double4> w(n);
    for (int64_t i = 0; i < n; i++)
      w[i] = make_double4(omega[3 * i], omega[3 * i + 1], omega[3 * i + 2],
                          (double)type_bits(type, inst_of, i));
    HIPCHK(h, hipMemcpy(h->st[0].omg, w.data(), sizeof(double4) * n, hipMemcpyHostToDevice));
  }
  HIPCHK(h, hipMemcpy(h->aux[0].tag, tag, sizeof(int) * n, hipMemcpyHostToDevice));
  HIPCHK(h, hipMemcpy(h->aux[0].orig, orig.data(), sizeof(int) * n, hipMemcpyHostToDevice));
  HIPCHK(h, hipMemcpy(h->aux[0].meta, zero.data(), sizeof(int) * n, hipMemcpyHostToDevice));
  HIPCHK(h, hipMemset(h->fout, 0, sizeof(double) * 6 * cap));
  h->cur_st = h->cur_aux = h->cur_l = 0;
  h->have_list = false;
  h->setup_done = false;
  return DEM_OK;
}

int dem_insert_particles(dem_handle *h, int64_t n, const double *x, const double *v,
                         const double *omega, const double *radius, const double *rmass,
                         const int *type, const int *tag) {
  if (!h) return DEM_ERR_ARG;
  if (n < 0 || (n > 0 && (!x || !v || !omega || !radius || !rmass || !type || !tag)))
    return fail(h, DEM_ERR_ARG, "bad particle arrays");
  if (!h->mat_set) return fail(h, DEM_ERR_STATE, "set materials before particles");
  if (h->ninst > 1) return fail(h, DEM_ERR_STATE, "insertion into batched instances is not available");
  for (int64_t i = 0; i < n; i++) {
    if (!(radius[i] > 0.0)) return fail(h, DEM_ERR_ARG, "Fix nve/sphere requires extended particles");
    if (!(rmass[i] > 0.0)) return fail(h, DEM_ERR_ARG, "particle mass must be > 0");
    if (type[i] < 1 || type[i] > h->ntypes) return fail(h, DEM_ERR_ARG, "Invalid atom type");
    if (tag[i] <= 0) return fail(h, DEM_ERR_ARG, "Invalid atom ID");
    for (int d = 0; d < 3; d++)
      if (!std::isfinite(x[3 * i + d]) || !std::isfinite(v[3 * i + d]) || !std::isfinite(omega[3 * i + d]))
        return fail(h, DEM_ERR_ARG, "non-finite particle state");
  }
  auto &q = h->pend;
  q.x.insert(q.x.end(), x, x + 3 * n);
  q.v.insert(q.v.end(), v, v + 3 * n);
  q.w.insert(q.w.end(), omega, omega + 3 * n);
  q.r.insert(q.r.end(), radius, radius + n);
  q.m.insert(q.m.end(), rmass, rmass + n);
  q.type.insert(q.type.end(), type, type + n);
  q.tag.insert(q.tag.end(), tag, tag + n);
  return DEM_OK;
}

int dem_set_max_radius(dem_handle *h, double rmax) {
  if (!h) return DEM_ERR_ARG;
  if (!(rmax >= 0.0) || !std::isfinite(rmax)) return fail(h, DEM_ERR_ARG, "bad radius");
  h->rmax_future = rmax;
  if (h->n > 0) {
    // the cells follow at the next build
    const double rm = std::max(h->rmax, h->rmax_future);
    if ((rm + rm) * h->p.cdf + h->p.skin != h->cutneighmax) {
      HIPCHK(h, hipStreamSynchronize(h->stream));
      for (int **p : {&h->cell_rank, &h->cell_count, &h->cell_start, &h->scan_tmp}) {
        dfree(h, *p);
        *p = nullptr;
      }
      int rc = setup_grid(h);
      if (rc) return rc;
      if (h->ntri > 0) {
        dfree(h, h->d_tcell_count);
        dfree(h, h->d_tcell_start);
        if ((rc = dalloc(h, (void **)&h->d_tcell_count, sizeof(int) * h->ncell))) return rc;
        if ((rc = dalloc(h, (void **)&h->d_tcell_start, sizeof(int) * (h->ncell + 1)))) return rc;
      }
    }
  }
  return DEM_OK;
}

int dem_setup(dem_handle *h) {
  if (!h) return DEM_ERR_ARG;
  return do_setup(h);
}

int dem_set_kernel_timing(dem_handle *h, int on) {
  if (!h) return DEM_ERR_ARG;
  h->timing = on != 0;
  h->kernel_ms = 0.0;
  h->kernel_launches = 0;
  h->rebuild_ms = 0.0;
  return DEM_OK;
}

int dem_set_step_shape(dem_handle *h, int variant, int persist) {
  if (!h) return DEM_ERR_ARG;
  if (variant < -1 || variant >= dem::kNumStepVariants || persist < 0 || persist > 64)
    return fail(h, DEM_ERR_ARG, "bad step kernel shape");
  clear_graphs(h);
  auto &tu = h->tuner;
  if (variant == -1) {
    tu = dem_handle::Tuner{};
    tu.active = true;
    const bool gen = uses_generic_models(h);
    for (int v = 0; v < dem::kNumStepVariants; v++) {
      if (gen && v != 0 && v != 5 && v != 7) continue;  // the sibling models' shapes
      for (int f : {0, 4, 8}) tu.cand.push_back({v, f});
    }
    tu.us.assign(tu.cand.size(), 0.0);
    return DEM_OK;
  }
  tu.active = false;
  h->variant = variant;
  h->persist = persist;
  return DEM_OK;
}

int dem_set_graph_capture(dem_handle *h, int on) {
  if (!h) return DEM_ERR_ARG;
  HIPCHK(h, hipStreamSynchronize(h->stream));
  clear_graphs(h);
  h->use_graphs = on != 0;
  return DEM_OK;
}

int dem_get_step_shape(dem_handle *h, int *variant, int *persist, int cap, int *ncand,
                       int *cand_variant, int *cand_persist, double *cand_us) {
  if (!h) return DEM_ERR_ARG;
  if (variant) *variant = h->variant;
  if (persist) *persist = h->persist;
  const auto &tu = h->tuner;
  if (ncand) *ncand = (int)tu.cand.size();
  for (int k = 0; k < (int)tu.cand.size() && k < cap; k++) {
    if (cand_variant) cand_variant[k] = tu.cand[k].first;
    if (cand_persist) cand_persist[k] = tu.cand[k].second;
    if (cand_us) cand_us[k] = tu.us[k];
  }
  return DEM_OK;
}

}  // extern "C"

namespace {

// Integrate nsteps steps continuing from the last force pass (whose f and
// torque are in fout, in the current particle order).
constexpr int kTuneBatch = 8;   // steps per candidate tile shape
constexpr int kGraphBatch = 32; // steps per captured graph

// The graph of kGraphBatch consecutive (non-final) steps starting with state
// buffer `st` and geometry buffer `tc`, captured on first use. Its launches
// take their step index from the control block (launch_set_base).
int step_graph(dem_handle *h, int st, int tc, bool moving, hipGraphExec_t *out) {
  const int key = st | (tc << 1) | (h->cur_aux << 2) | (h->variant << 3) | (h->persist << 6) |
                  ((moving ? 1 : 0) << 13);
  for (auto &g : h->graphs)
    if (g.key == key) {
      *out = g.exec;
      return DEM_OK;
    }
  HIPCHK(h, hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
  hipError_t err = hipSuccess;
  for (int b = 0; b < kGraphBatch && err == hipSuccess; b++) {
    dem::StepArgs a = step_args(h, (st + b) & 1, (tc + b) & 1);
    a.step = b;
    a.step_from_ctrl = 1;
    a.shearupdate = 1;
    a.do_final = 1;
    a.do_initial = 1;
    a.store_f = 0;
    a.count = 0;
    a.move_mesh = moving ? 1 : 0;
    err = dem::launch_step(a, h->stream);
  }
  hipGraph_t graph = nullptr;
  const hipError_t e2 = hipStreamEndCapture(h->stream, &graph);
  if (err != hipSuccess) return fail(h, DEM_ERR_HIP, hipGetErrorString(err));
  if (e2 != hipSuccess) return fail(h, DEM_ERR_HIP, hipGetErrorString(e2));
  hipGraphExec_t exec = nullptr;
  const hipError_t e3 = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  hipGraphDestroy(graph);
  if (e3 != hipSuccess) return fail(h, DEM_ERR_HIP, hipGetErrorString(e3));
  h->graphs.push_back({key, exec});
  *out = exec;
  return DEM_OK;
}

// passes over the candidate tile shapes: each keeps its fastest batch, so a
// shape measured while the state changed (contacts forming, a rebuild near)
// is measured again later
constexpr int kTuneRounds = 3;

// picks the fastest measured tile shape
void finish_tuning(dem_handle *h) {
  auto &tu = h->tuner;
  int best = -1;
  for (int k = 0; k < (int)tu.cand.size(); k++)
    if (tu.us[k] > 0.0 && (best < 0 || tu.us[k] < tu.us[best])) best = k;
  clear_graphs(h);
  if (best >= 0) {
    h->variant = tu.cand[best].first;
    h->persist = tu.cand[best].second;
  } else {
    h->variant = 0;
    h->persist = 0;
  }
  tu.active = false;
}

int run_steps_piece(dem_handle *h, int64_t nsteps);

// A run with time-dependent movers goes in pieces of at most kMotionSteps
// steps (bitwise the same as one run: the pieces continue as `run pre no`),
// so their motion table stays a bounded size.
int run_steps(dem_handle *h, int64_t nsteps) {
  if (!has_timed_movers(h)) return run_steps_piece(h, nsteps);
  for (int64_t done = 0; done < nsteps;) {
    const int64_t k = std::min<int64_t>(kMotionSteps, nsteps - done);
    int rc = run_steps_piece(h, k);
    if (rc) return rc;
    done += k;
  }
  return DEM_OK;
}

int run_steps_piece(dem_handle *h, int64_t nsteps) {
  int rc;
  if (nsteps == 0) return DEM_OK;
  // particles of dem_insert_particles join at this run's first step
  const bool joining = !h->pend.tag.empty();
  if (h->n == 0) {
    if (any_moving_mesh(h)) return fail(h, DEM_ERR_STATE, "moving meshes need particles on this path");
    if (!joining) {  // nothing to integrate: the steps pass
      h->ntimestep += nsteps;
      h->can_continue = true;
      return DEM_OK;
    }
  }
  h->setup_done = false;
  h->can_continue = false;
  // a rebuild a moving mesh asked for at the first step of this run
  // (FixMesh::pre_force set next_reneighbor on the last step of the previous
  // run), or the build an insertion forces (fix_insert.cpp:311-319)
  HIPCHK(h, dem::launch_set_stop(h->ctrl, (h->pending_mesh_rebuild || joining) ? 1 : kNever, h->stream,
                                      h->d_inst_stats, 2 * h->ninst));
  h->pending_mesh_rebuild = false;
  const bool moving = any_moving_mesh(h);
  const int64_t ts0 = h->ntimestep;  // step t of this run is timestep ts0 + t
  const bool timed = has_timed_movers(h);
  double t_end = h->mesh_time;
  if (timed && (rc = upload_mesh_motion(h, nsteps, &t_end))) return rc;
  if (h->n > 0) {
    dem::StepArgs a = step_args(h, h->cur_st, h->tri_cur);
    a.ago_next = ts0 + 1 - h->last_build;
    HIPCHK(h, dem::launch_initial(a, h->stream));
  }
  // the inserted particles: after the initial integrate of the others, before
  // the forced build (Verlet::run: initial_integrate, then pre_exchange on a
  // build step, verlet.cpp:283-320)
  if (joining) {
    if ((rc = append_pending(h))) return rc;
    HIPCHK(h, dem::launch_set_stop(h->ctrl, 1, h->stream, h->d_inst_stats, 2 * h->ninst));
  }
  // the first step's mesh rotation, in place; every later step's rotation is
  // done by the trailing blocks of the step before
  if (moving) HIPCHK(h, dem::launch_mesh_move(mesh_move_args(h, 1, 0, 1), h->stream));
  int tc = h->tri_cur;  // geometry buffer of step t
  int64_t t = 1;
  int batch = 16;
  auto &tu = h->tuner;
  while (t <= nsteps) {
    // while tuning, each batch runs one candidate tile shape; every shape
    // computes the same result, so these are ordinary steps of the run
    const bool tuning = tu.active && tu.idx < (int)tu.cand.size();
    if (tuning) {
      h->variant = tu.cand[tu.idx].first;
      h->persist = tu.cand[tu.idx].second;
    }
    const bool events = h->timing || tuning;
    int nb = (int)std::min<int64_t>(tuning ? kTuneBatch : batch, nsteps - t + 1);
    // whole graphs of kGraphBatch steps, none of them the run's last step
    const bool graphed = h->use_graphs && !events && nb >= kGraphBatch && t + nb - 1 < nsteps;
    if (graphed) {
      nb -= nb % kGraphBatch;
      for (int g0 = 0; g0 < nb; g0 += kGraphBatch) {
        const int64_t step0 = t + g0;
        HIPCHK(h, dem::launch_set_base(h->ctrl, step0, ts0 + step0 + 1 - h->last_build, h->stream));
        hipGraphExec_t exec = nullptr;
        if ((rc = step_graph(h, (h->cur_st + g0) & 1, (tc + g0) & 1, moving, &exec))) return rc;
        HIPCHK(h, hipGraphLaunch(exec, h->stream));
      }
    }
    for (int b = 0; b < nb && !graphed; b++) {
      const int64_t step = t + b;
      const bool last = step == nsteps;
      dem::StepArgs a = step_args(h, (h->cur_st + b) & 1, (tc + b) & 1);
      a.step = step;
      a.shearupdate = 1;
      a.do_final = 1;
      a.do_initial = last ? 0 : 1;
      a.store_f = last ? 1 : 0;
      a.count = last ? 1 : 0;
      a.move_mesh = (moving && !last) ? 1 : 0;
      a.mv.step = step + 1;
      a.ago_next = ts0 + step + 1 - h->last_build;
      HIPCHK(h, dem::launch_step(a, h->stream));
      // the start event goes after the batch's first launch: the GPU is then
      // busy while the host enqueues the rest, so the pair times the kernels
      // and not the latency of the first launch
      if (events && b == 0) HIPCHK(h, hipEventRecord(h->ev[0], h->stream));
    }
    if (events) HIPCHK(h, hipEventRecord(h->ev[1], h->stream));
    if ((rc = read_ctrl(h))) return rc;
    if (h->ctrl_host->slot_overflow)
      return fail(h, DEM_ERR_OVERFLOW, "more simultaneous mesh contacts on a particle than history slots");
    const long long stop = h->ctrl_host->stop_at;
    // a stop beyond this batch (a mesh asks for a rebuild two steps after the
    // step that moved it) stays in the control block for the next batch
    const bool rebuild_now = stop <= nsteps && stop - t <= nb;
    const int executed = rebuild_now ? (int)(stop - t) : nb;
    if (executed < 0 || executed > nb) return fail(h, DEM_ERR_STATE, "internal step accounting error");
    // Timing brackets launches 1..nb-1 of the batch with one event pair
    // (events between launches would add their own gaps); batches that ended
    // in a rebuild request contain skipped launches and are left out.
    if (events && executed == nb && nb > 1) {
      float ms = 0.f;
      HIPCHK(h, hipEventElapsedTime(&ms, h->ev[0], h->ev[1]));
      if (h->timing) {
        h->kernel_ms += ms;
        h->kernel_launches += nb - 1;
      }
      if (tuning) {
        const double us = 1e3 * ms / (nb - 1);
        if (tu.us[tu.idx] <= 0.0 || us < tu.us[tu.idx]) tu.us[tu.idx] = us;
        tu.idx++;
        tu.tries = 0;
      }
    } else if (tuning && ++tu.tries >= 3) {
      tu.idx++;  // interrupted by rebuilds three times: leave this shape out of the pass
      tu.tries = 0;
    }
    if (tuning && tu.idx == (int)tu.cand.size()) {
      if (++tu.round < kTuneRounds)
        tu.idx = 0;
      else
        finish_tuning(h);
    }
    const bool ran_last = t + executed - 1 == nsteps;  // the last step moves no mesh
    h->cur_st = (h->cur_st + executed) & 1;
    h->ntimestep += executed;
    tc = (tc + executed - (ran_last && moving ? 1 : 0)) & 1;
    if (!moving) tc = h->tri_cur;
    t += executed;
    if (rebuild_now) {
      h->tri_cur = tc;
      if ((rc = rebuild(h))) return rc;
      h->last_build = ts0 + stop;  // the build belongs to step `stop` (its initial integrate is done)
      HIPCHK(h, dem::launch_set_stop(h->ctrl, kNever, h->stream, h->d_inst_stats, 2 * h->ninst));
      batch = std::max(4, batch / 2);
    } else {
      batch = std::min(256, batch * 2);
    }
  }
  h->tri_cur = tc;
  if (timed) h->mesh_time = t_end;
  h->pending_mesh_rebuild = h->ctrl_host->stop_at == nsteps + 1;
  h->ncontacts = (int64_t)(h->ctrl_host->ncontact_full / 2);
  h->nwall_contacts = (int64_t)h->ctrl_host->nwall;
  h->can_continue = true;
  return DEM_OK;
}

}  // namespace

extern "C" {

int dem_run(dem_handle *h, int64_t nsteps) {
  if (!h) return DEM_ERR_ARG;
  if (nsteps < 0) return fail(h, DEM_ERR_ARG, "negative step count");
  int rc = do_setup(h);  // Run::command calls Integrate::setup (run.cpp:209)
  if (rc) return rc;
  h->can_continue = true;
  return run_steps(h, nsteps);
}

int dem_run_continue(dem_handle *h, int64_t nsteps) {
  if (!h) return DEM_ERR_ARG;
  if (nsteps < 0) return fail(h, DEM_ERR_ARG, "negative step count");
  if (!h->can_continue)
    return fail(h, DEM_ERR_STATE, "run pre no needs a preceding run or setup on this state");
  return run_steps(h, nsteps);
}

int dem_compute_forces(dem_handle *h, int shearupdate) {
  int rc = check_ready(h);
  if (rc) return rc;
  if (!h->have_list) {
    if ((rc = rebuild(h))) return rc;
  }
  return force_pass(h, shearupdate ? 1 : 0);
}

int dem_get_box(dem_handle *h, double *boxlo, double *boxhi) {
  if (!h || !boxlo || !boxhi) return DEM_ERR_ARG;
  const dem_params &p = h->p;
  for (int d = 0; d < 3; d++) {
    boxlo[d] = p.boxlo[d];
    boxhi[d] = p.boxhi[d];
  }
  if (!p.shrink_faces) return DEM_OK;
  if (h->n <= 0 || !h->have_list) return DEM_OK;
  // Domain::reset_box (domain.cpp:381-449): the extent of the atoms where the
  // last build saw them (xhold) plus small = 1e-4 of the created box length
  // (domain.cpp:85, 254-259); an m face not inside its created position
  HIPCHK(h, hipStreamSynchronize(h->stream));
  const size_t cap = (size_t)h->cap;
  std::vector<double> xh(3 * cap);
  HIPCHK(h, hipMemcpy(xh.data(), h->xhold, sizeof(double) * 3 * cap, hipMemcpyDeviceToHost));
  for (int d = 0; d < 3; d++) {
    double mn = 1.0e20, mx = -1.0e20;  // BIG of domain.cpp:84
    for (int64_t i = 0; i < h->n; i++) {
      mn = std::min(mn, xh[d * cap + i]);
      mx = std::max(mx, xh[d * cap + i]);
    }
    const double small = 1.0e-4 * (p.boxhi[d] - p.boxlo[d]);
    if ((p.shrink_faces >> (2 * d)) & 1) {
      boxlo[d] = -(-mn) - small;
      if ((p.shrink_min >> (2 * d)) & 1) boxlo[d] = std::min(boxlo[d], p.boxlo[d]);
    }
    if ((p.shrink_faces >> (2 * d + 1)) & 1) {
      boxhi[d] = mx + small;
      if ((p.shrink_min >> (2 * d + 1)) & 1) boxhi[d] = std::max(boxhi[d], p.boxhi[d]);
    }
  }
  return DEM_OK;
}

int dem_get_state(dem_handle *h, double *x, double *v, double *omega, double *f, double *torque) {
  if (!h) return DEM_ERR_ARG;
  if (h->n <= 0) return DEM_OK;  // an empty system (particles still to be inserted): nothing to copy
  HIPCHK(h, hipStreamSynchronize(h->stream));
  const int64_t n = h->n;
  const size_t cap = (size_t)h->cap;
  std::vector<double4> pos(n), vel(n);
  std::vector<int> orig(n);
  const StateBuf &s = h->st[h->cur_st];
  HIPCHK(h, hipMemcpy(pos.data(), s.pos, sizeof(double4) * n, hipMemcpyDeviceToHost));
  HIPCHK(h, hipMemcpy(vel.data(), s.vel, sizeof(double4) * n, hipMemcpyDeviceToHost));
  HIPCHK(h, hipMemcpy(orig.data(), h->aux[h->cur_aux].orig, sizeof(int) * n, hipMemcpyDeviceToHost));
  std::vector<double> w(3 * n);
  if (h->mixed) {
    std::vector<float4> ww(n);
    HIPCHK(h, hipMemcpy(ww.data(), s.omg, sizeof(float4) * n, hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < n; i++) {
      w[3 * i] = ww[i].x;
      w[3 * i + 1] = ww[i].y;
      w[3 * i + 2] = ww[i].z;
    }
  } else {
    std::vector<double4> ww(n);
    HIPCHK(h, hipMemcpy(ww.data(), s.omg, sizeof(double4) * n, hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < n; i++) {
      w[3 * i] = ww[i].x;
      w[3 * i + 1] = ww[i].y;
      w[3 * i + 2] = ww[i].z;
    }
  }
  std::vector<double> ft;
  if (f || torque) {
    ft.resize(6 * cap);
    HIPCHK(h, hipMemcpy(ft.data(), h->fout, sizeof(double) * 6 * cap, hipMemcpyDeviceToHost));
  }
  for (int64_t i = 0; i < n; i++) {
    const int o = orig[i];
    if (o < 0 || o >= n) return fail(h, DEM_ERR_STATE, "corrupt particle index");
    if (x) {
      x[3 * o] = pos[i].x;
      x[3 * o + 1] = pos[i].y;
      x[3 * o + 2] = pos[i].z;
    }
    if (v) {
      v[3 * o] = vel[i].x;
      v[3 * o + 1] = vel[i].y;
      v[3 * o + 2] = vel[i].z;
    }
    if (omega)
      for (int d = 0; d < 3; d++) omega[3 * o + d] = w[3 * i + d];
    if (f)
      for (int d = 0; d < 3; d++) f[3 * o + d] = ft[d * cap + i];
    if (torque)
      for (int d = 0; d < 3; d++) torque[3 * o + d] = ft[(3 + d) * cap + i];
  }
  return DEM_OK;
}

// ---------------------------------------------------------------------------
// restart files: the particles in caller order (what read_restart restores
// into atom arrays) followed by a snapshot of the device state in its own
// order (state records, neighbor list with contact flags and histories, wall
// and mesh histories, displacement reference, mesh geometry) and the counters
// a run continues from. The layout is this library's own.
// ---------------------------------------------------------------------------
}  // extern "C"

namespace {
struct RestartHeader {
  char magic[8];
  int32_t version, mixed, ntypes, cap, W, nwalls, nmesh, ntri;
  int32_t cur_st, cur_l, cur_aux, tri_cur, have_list, pending_mesh_rebuild, periodic[3];
  int32_t hist_w;  // history values per pair / plane contact (files without it: 0, meaning 3)
  int64_t n, ntimestep, nrebuilds, last_build;
  double mesh_time, dt, boxlo[3], boxhi[3];
};
constexpr char kRestartMagic[8] = {'L', 'G', 'A', 'M', 'D', 'R', 'S', 'T'};
// 2: pair histories stored as one record per list entry
constexpr int32_t kRestartVersion = 2;

bool read_header(FILE *f, RestartHeader &hd) {
  return fread(&hd, sizeof hd, 1, f) == 1 && std::memcmp(hd.magic, kRestartMagic, 8) == 0 &&
         hd.version == kRestartVersion;
}

// device blobs of a restart, in file order
std::vector<std::pair<void *, size_t>> restart_blobs(dem_handle *h) {
  const size_t cap = (size_t)h->cap, n = (size_t)h->n;
  const size_t nw = std::max<size_t>(1, h->walls.size()), nm = std::max<size_t>(1, h->meshes.size());
  const StateBuf &st = h->st[h->cur_st];
  const AuxBuf &ax = h->aux[h->cur_aux];
  const ListBuf &ls = h->lst[h->cur_l];
  std::vector<std::pair<void *, size_t>> b = {
      {st.pos, sizeof(double4) * n},
      {st.vel, sizeof(double4) * n},
      {st.omg, (h->mixed ? sizeof(float4) : sizeof(double4)) * n},
      {ax.tag, sizeof(int) * n},
      {ax.orig, sizeof(int) * n},
      {ax.meta, sizeof(int) * n},
      {ax.whist, h->fsz * h->hw * nw * cap},
      {ax.mslot, sizeof(int) * nm * dem::kMeshSlots * cap},
      {ax.mhist, h->fsz * h->hw * nm * dem::kMeshSlots * cap},
      {ls.nbr, sizeof(uint32_t) * h->W * cap},
      {ls.hist, h->fsz * dem::hist_rec(h->hw) * h->W * cap},
      {h->xhold, sizeof(double) * 3 * cap},
      {h->fout, sizeof(double) * 6 * cap}};
  if (h->ntri > 0) {
    b.push_back({h->d_tri_buf[h->tri_cur], sizeof(dem::TriGeom) * h->ntri});
    b.push_back({h->d_node_last, sizeof(double) * 9 * h->ntri});
  }
  return b;
}
}  // namespace

extern "C" {

int dem_write_restart(dem_handle *h, const char *path) {
  if (!h || !path) return DEM_ERR_ARG;
  if (h->n <= 0 || !h->have_list) return fail(h, DEM_ERR_STATE, "write_restart needs a set-up system");
  const int64_t n = h->n;
  std::vector<double> x(3 * n), v(3 * n), w(3 * n);
  int rc = dem_get_state(h, x.data(), v.data(), w.data(), nullptr, nullptr);
  if (rc) return rc;
  // caller-order tag, type, radius, mass
  std::vector<double4> pos(n), vel(n);
  std::vector<int> orig(n), tag(n), typ(n), ctag(n), ctype(n);
  std::vector<double> crad(n), cmass(n);
  const StateBuf &st = h->st[h->cur_st];
  const AuxBuf &ax = h->aux[h->cur_aux];
  HIPCHK(h, hipMemcpy(pos.data(), st.pos, sizeof(double4) * n, hipMemcpyDeviceToHost));
  HIPCHK(h, hipMemcpy(vel.data(), st.vel, sizeof(double4) * n, hipMemcpyDeviceToHost));
  HIPCHK(h, hipMemcpy(orig.data(), ax.orig, sizeof(int) * n, hipMemcpyDeviceToHost));
  HIPCHK(h, hipMemcpy(tag.data(), ax.tag, sizeof(int) * n, hipMemcpyDeviceToHost));
  if (h->mixed) {
    std::vector<float4> ww(n);
    HIPCHK(h, hipMemcpy(ww.data(), st.omg, sizeof(float4) * n, hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < n; i++) typ[i] = (int)ww[i].w & dem::kTypeMask;
  } else {
    std::vector<double4> ww(n);
    HIPCHK(h, hipMemcpy(ww.data(), st.omg, sizeof(double4) * n, hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < n; i++) typ[i] = (int)ww[i].w & dem::kTypeMask;
  }
  for (int64_t i = 0; i < n; i++) {
    const int o = orig[i];
    ctag[o] = tag[i];
    ctype[o] = typ[i];
    crad[o] = pos[i].w;
    cmass[o] = vel[i].w;
  }
  FILE *f = fopen(path, "wb");
  if (!f) return fail(h, DEM_ERR_ARG, std::string("Cannot open restart file ") + path);
  RestartHeader hd{};
  std::memcpy(hd.magic, kRestartMagic, 8);
  hd.version = kRestartVersion;
  hd.mixed = h->mixed ? 1 : 0;
  hd.ntypes = h->ntypes;
  hd.cap = h->cap;
  hd.W = h->W;
  hd.hist_w = h->hw;
  hd.nwalls = (int32_t)h->walls.size();
  hd.nmesh = (int32_t)h->meshes.size();
  hd.ntri = h->ntri;
  hd.cur_st = h->cur_st;
  hd.cur_l = h->cur_l;
  hd.cur_aux = h->cur_aux;
  hd.tri_cur = h->tri_cur;
  hd.have_list = h->have_list ? 1 : 0;
  hd.pending_mesh_rebuild = h->pending_mesh_rebuild ? 1 : 0;
  hd.n = n;
  hd.ntimestep = h->ntimestep;
  hd.nrebuilds = h->nrebuilds;
  hd.last_build = h->last_build;
  hd.mesh_time = h->mesh_time;
  hd.dt = h->p.dt;
  for (int d = 0; d < 3; d++) {
    hd.boxlo[d] = h->p.boxlo[d];
    hd.boxhi[d] = h->p.boxhi[d];
    hd.periodic[d] = h->p.periodic[d];
  }
  bool ok = fwrite(&hd, sizeof hd, 1, f) == 1;
  ok = ok && fwrite(ctag.data(), sizeof(int), n, f) == (size_t)n;
  ok = ok && fwrite(ctype.data(), sizeof(int), n, f) == (size_t)n;
  ok = ok && fwrite(x.data(), sizeof(double), 3 * n, f) == (size_t)(3 * n);
  ok = ok && fwrite(v.data(), sizeof(double), 3 * n, f) == (size_t)(3 * n);
  ok = ok && fwrite(w.data(), sizeof(double), 3 * n, f) == (size_t)(3 * n);
  ok = ok && fwrite(crad.data(), sizeof(double), n, f) == (size_t)n;
  ok = ok && fwrite(cmass.data(), sizeof(double), n, f) == (size_t)n;
  for (auto &blob : restart_blobs(h)) {
    if (!ok) break;
    std::vector<char> buf(blob.second);
    if (hipMemcpy(buf.data(), blob.first, blob.second, hipMemcpyDeviceToHost) != hipSuccess) {
      fclose(f);
      return fail(h, DEM_ERR_HIP, "restart: device read failed");
    }
    ok = fwrite(buf.data(), 1, buf.size(), f) == buf.size();
  }
  if (fclose(f) != 0) ok = false;
  if (!ok) return fail(h, DEM_ERR_ARG, std::string("Error writing restart file ") + path);
  return DEM_OK;
}

int dem_restart_info(const char *path, int64_t *n, int *ntypes, double *boxlo, double *boxhi,
                     int *periodic, int64_t *ntimestep) {
  if (!path) return DEM_ERR_ARG;
  FILE *f = fopen(path, "rb");
  if (!f) return fail(nullptr, DEM_ERR_ARG, std::string("Cannot open restart file ") + path);
  RestartHeader hd;
  const bool ok = read_header(f, hd);
  fclose(f);
  if (!ok) return fail(nullptr, DEM_ERR_ARG, std::string("Not a restart file of this library: ") + path);
  if (n) *n = hd.n;
  if (ntypes) *ntypes = hd.ntypes;
  for (int d = 0; d < 3; d++) {
    if (boxlo) boxlo[d] = hd.boxlo[d];
    if (boxhi) boxhi[d] = hd.boxhi[d];
    if (periodic) periodic[d] = hd.periodic[d];
  }
  if (ntimestep) *ntimestep = hd.ntimestep;
  return DEM_OK;
}

int dem_restart_particles(const char *path, int64_t n, int *tag, int *type, double *x, double *v,
                          double *omega, double *radius, double *rmass) {
  if (!path) return DEM_ERR_ARG;
  FILE *f = fopen(path, "rb");
  if (!f) return fail(nullptr, DEM_ERR_ARG, std::string("Cannot open restart file ") + path);
  RestartHeader hd;
  bool ok = read_header(f, hd) && hd.n == n;
  std::vector<int> ib(n);
  std::vector<double> db(3 * n);
  auto rd = [&](void *dst, size_t sz, size_t cnt) {
    if (!ok) return;
    void *tgt = dst ? dst : (sz == sizeof(int) ? (void *)ib.data() : (void *)db.data());
    ok = fread(tgt, sz, cnt, f) == cnt;
  };
  rd(tag, sizeof(int), n);
  rd(type, sizeof(int), n);
  rd(x, sizeof(double), 3 * n);
  rd(v, sizeof(double), 3 * n);
  rd(omega, sizeof(double), 3 * n);
  rd(radius, sizeof(double), n);
  rd(rmass, sizeof(double), n);
  fclose(f);
  if (!ok) return fail(nullptr, DEM_ERR_ARG, std::string("Bad or mismatched restart file ") + path);
  return DEM_OK;
}

int dem_read_restart(dem_handle *h, const char *path) {
  if (!h || !path) return DEM_ERR_ARG;
  if (h->n <= 0) return fail(h, DEM_ERR_STATE, "read_restart: set the particles first");
  FILE *f = fopen(path, "rb");
  if (!f) return fail(h, DEM_ERR_ARG, std::string("Cannot open restart file ") + path);
  RestartHeader hd;
  if (!read_header(f, hd)) {
    fclose(f);
    return fail(h, DEM_ERR_ARG, std::string("Not a restart file of this library: ") + path);
  }
  if ((hd.hist_w ? hd.hist_w : 3) != h->hw) {
    fclose(f);
    return fail(h, DEM_ERR_ARG,
                "read_restart: the file was written with a different rolling model (epsd2 history)");
  }
  if (hd.n != h->n || hd.cap != h->cap || hd.W != h->W || hd.mixed != (h->mixed ? 1 : 0) ||
      hd.nwalls != (int)h->walls.size() || hd.nmesh != (int)h->meshes.size() || hd.ntri != h->ntri ||
      hd.dt != h->p.dt) {
    fclose(f);
    return fail(h, DEM_ERR_ARG,
                "read_restart: the system differs from the one written (particles, list width, "
                "precision, walls, meshes or timestep)");
  }
  // skip the caller-order particle table
  const long skip = (long)(hd.n * (2 * sizeof(int) + 11 * sizeof(double)));
  if (fseek(f, skip, SEEK_CUR) != 0) {
    fclose(f);
    return fail(h, DEM_ERR_ARG, "read_restart: truncated file");
  }
  HIPCHK(h, hipStreamSynchronize(h->stream));
  clear_graphs(h);
  h->cur_st = hd.cur_st;
  h->cur_l = hd.cur_l;
  h->cur_aux = hd.cur_aux;
  h->tri_cur = hd.tri_cur;
  bool ok = true;
  for (auto &blob : restart_blobs(h)) {
    std::vector<char> buf(blob.second);
    if (fread(buf.data(), 1, buf.size(), f) != buf.size()) {
      ok = false;
      break;
    }
    if (hipMemcpy(blob.first, buf.data(), blob.second, hipMemcpyHostToDevice) != hipSuccess) {
      fclose(f);
      return fail(h, DEM_ERR_HIP, "read_restart: device write failed");
    }
  }
  fclose(f);
  if (!ok) return fail(h, DEM_ERR_ARG, "read_restart: truncated file");
  h->have_list = hd.have_list != 0;
  h->pending_mesh_rebuild = hd.pending_mesh_rebuild != 0;
  h->ntimestep = hd.ntimestep;
  h->nrebuilds = hd.nrebuilds;
  h->last_build = hd.last_build;
  h->mesh_time = hd.mesh_time;
  h->setup_done = false;
  h->can_continue = false;  // the next run sets up (Verlet::setup), as after read_restart
  return DEM_OK;
}

}  // extern "C"

namespace {
// pairs of the half list with history values comp0..comp0+2 (0 shear, 3 the
// epsd2 rolling torque)
int get_pairs(dem_handle *h, int64_t cap_out, int *tags, int *flags, double *hist,
              int64_t *npairs, int comp0) {
  if (!h || !npairs) return DEM_ERR_ARG;
  *npairs = 0;
  if (!h->have_list) return DEM_OK;
  HIPCHK(h, hipStreamSynchronize(h->stream));
  const int64_t n = h->n;
  const size_t cap = (size_t)h->cap;
  const int W = h->W;
  std::vector<uint32_t> nbr((size_t)W * cap);
  std::vector<int> meta(n), tag(n);
  const size_t rec = dem::hist_rec(h->hw);
  std::vector<char> hraw(h->fsz * rec * W * cap);
  const ListBuf &l = h->lst[h->cur_l];
  HIPCHK(h, hipMemcpy(nbr.data(), l.nbr, sizeof(uint32_t) * W * cap, hipMemcpyDeviceToHost));
  HIPCHK(h, hipMemcpy(hraw.data(), l.hist, hraw.size(), hipMemcpyDeviceToHost));
  HIPCHK(h, hipMemcpy(meta.data(), h->aux[h->cur_aux].meta, sizeof(int) * n, hipMemcpyDeviceToHost));
  HIPCHK(h, hipMemcpy(tag.data(), h->aux[h->cur_aux].tag, sizeof(int) * n, hipMemcpyDeviceToHost));
  int64_t np = 0;
  for (int64_t i = 0; i < n; i++) {
    const int cnt = meta[i] & dem::kMetaCountMask;
    for (int k = 0; k < cnt; k++) {
      const uint32_t e = nbr[(size_t)k * cap + i];
      const int j = (int)(e & dem::kIndexMask);
      if (j < 0 || j >= n) return fail(h, DEM_ERR_STATE, "corrupt neighbor entry");
      if (!(tag[i] < tag[j])) continue;
      if (np < cap_out) {
        tags[2 * np] = tag[i];
        tags[2 * np + 1] = tag[j];
        const bool fl = (e & dem::kFlagBit) != 0;
        flags[np] = fl ? 1 : 0;
        for (int d = 0; d < 3; d++) {
          const size_t idx = dem::hist_at(k, i, cap, rec) + dem::hist_off(comp0 + d);
          double val = h->mixed ? (double)((float *)hraw.data())[idx] : ((double *)hraw.data())[idx];
          hist[3 * np + d] = fl ? val : 0.0;
        }
      }
      np++;
    }
  }
  *npairs = np;
  return DEM_OK;
}

int get_wall_history(dem_handle *h, double *out, int comp0) {
  if (!h || !out) return DEM_ERR_ARG;
  if (h->n <= 0) return fail(h, DEM_ERR_STATE, "no particles");
  HIPCHK(h, hipStreamSynchronize(h->stream));
  const int64_t n = h->n;
  const size_t cap = (size_t)h->cap;
  const int nw = (int)h->walls.size();
  if (nw == 0) return DEM_OK;
  std::vector<int> meta(n), orig(n);
  std::vector<char> raw(h->fsz * h->hw * nw * cap);
  const AuxBuf &a = h->aux[h->cur_aux];
  HIPCHK(h, hipMemcpy(meta.data(), a.meta, sizeof(int) * n, hipMemcpyDeviceToHost));
  HIPCHK(h, hipMemcpy(orig.data(), a.orig, sizeof(int) * n, hipMemcpyDeviceToHost));
  HIPCHK(h, hipMemcpy(raw.data(), a.whist, raw.size(), hipMemcpyDeviceToHost));
  for (int64_t i = 0; i < n; i++) {
    const int o = orig[i];
    for (int w = 0; w < nw; w++) {
      const bool has = meta[i] & (1 << (dem::kMetaWallShift + w));
      for (int d = 0; d < 3; d++) {
        const size_t idx = ((size_t)w * h->hw + comp0 + d) * cap + i;
        const double val = h->mixed ? (double)((float *)raw.data())[idx] : ((double *)raw.data())[idx];
        out[((size_t)o * nw + w) * 3 + d] = has ? val : 0.0;
      }
    }
  }
  return DEM_OK;
}
}  // namespace

extern "C" {

int dem_get_pairs(dem_handle *h, int64_t cap_out, int *tags, int *flags, double *hist,
                  int64_t *npairs) {
  return get_pairs(h, cap_out, tags, flags, hist, npairs, 0);
}

int dem_get_wall_history(dem_handle *h, double *out) { return get_wall_history(h, out, 0); }

static_assert(dem::kPairLocalFixedCols == DEM_PAIR_LOCAL_FIXED_COLS, "pair-local record layout and the header disagree");

int dem_compute_pair_local(dem_handle *h, int64_t cap_out, int *tags, double *data, int64_t *npairs,
                           int *ncol) {
  if (!h || !npairs) return DEM_ERR_ARG;
  *npairs = 0;
  const int nc = dem::kPairLocalFixedCols + h->hw;
  if (ncol) *ncol = nc;
  if (h->n <= 0 || !h->have_list) return DEM_OK;
  if (cap_out > 0 && (!tags || !data)) return fail(h, DEM_ERR_ARG, "null output arrays");
  HIPCHK(h, hipSetDevice(h->p.device));
  int rc = check_ready(h);
  if (rc) return rc;
  const int n = (int)h->n;
  std::vector<void *> tmp;
  auto tfree = [&]() {
    for (void *q : tmp) hipFree(q);
  };
  auto talloc = [&](void **q, size_t bytes) -> int {
    if (hipMalloc(q, bytes ? bytes : 16) != hipSuccess) return fail(h, DEM_ERR_HIP, "hipMalloc failed");
    tmp.push_back(*q);
    return DEM_OK;
  };
  dem::PairLocalArgs o{};
  o.ncol = nc;
  int *scan = nullptr;
  if ((rc = talloc((void **)&o.count, sizeof(int) * n)) || (rc = talloc((void **)&o.offset, sizeof(int) * (n + 1))) ||
      (rc = talloc((void **)&scan, sizeof(int) * dem::scan_tmp_size(n)))) {
    tfree();
    return rc;
  }
  dem::StepArgs a = step_args(h, h->cur_st, h->tri_cur);
  a.shearupdate = 0;
  hipError_t e = dem::launch_pair_local_count(a, o, h->stream);
  if (e == hipSuccess) e = dem::launch_exclusive_scan(o.count, o.offset, scan, n, h->stream);
  int total = 0;
  if (e == hipSuccess) e = hipMemcpyAsync(&total, o.offset + n, sizeof(int), hipMemcpyDeviceToHost, h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  if (e != hipSuccess) {
    tfree();
    return fail(h, DEM_ERR_HIP, hipGetErrorString(e));
  }
  *npairs = total;
  if (cap_out <= 0 || total == 0) {
    tfree();
    return DEM_OK;
  }
  if ((rc = talloc((void **)&o.tags, sizeof(int) * 2 * (size_t)total)) ||
      (rc = talloc((void **)&o.rec, sizeof(double) * (size_t)nc * (size_t)total))) {
    tfree();
    return rc;
  }
  e = dem::launch_pair_local_fill(a, o, h->stream);
  const int64_t m = std::min<int64_t>(cap_out, total);
  if (e == hipSuccess) e = hipMemcpyAsync(tags, o.tags, sizeof(int) * 2 * (size_t)m, hipMemcpyDeviceToHost, h->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(data, o.rec, sizeof(double) * (size_t)nc * (size_t)m, hipMemcpyDeviceToHost, h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  tfree();
  if (e != hipSuccess) return fail(h, DEM_ERR_HIP, hipGetErrorString(e));
  // Domain::remap of both positions (compute_pair_gran_local.cpp:448-451)
  for (int64_t q = 0; q < m; q++)
    for (int side = 0; side < 2; side++)
      for (int d = 0; d < 3; d++) {
        if (!h->p.periodic[d]) continue;
        double &x = data[q * nc + 3 * side + d];
        const double lo = h->p.boxlo[d], L = h->p.boxhi[d] - h->p.boxlo[d];
        while (x < lo) x += L;
        while (x >= lo + L) x -= L;
      }
  return DEM_OK;
}

static_assert(dem::kWallLocalFixedCols == DEM_WALL_LOCAL_FIXED_COLS, "wall-local record layout and the header disagree");

int dem_compute_wall_local(dem_handle *h, int64_t cap_out, int *ids, double *data, int64_t *nrows, int *ncol) {
  if (!h || !nrows) return DEM_ERR_ARG;
  *nrows = 0;
  const int nc = dem::kWallLocalFixedCols + 3;
  if (ncol) *ncol = nc;
  if (h->n <= 0 || !h->have_list || h->meshes.empty()) return DEM_OK;
  if (cap_out > 0 && (!ids || !data)) return fail(h, DEM_ERR_ARG, "null output arrays");
  HIPCHK(h, hipSetDevice(h->p.device));
  int rc = check_ready(h);
  if (rc) return rc;
  const int n = (int)h->n;
  std::vector<void *> tmp;
  auto tfree = [&]() {
    for (void *q : tmp) hipFree(q);
  };
  auto talloc = [&](void **q, size_t bytes) -> int {
    if (hipMalloc(q, bytes ? bytes : 16) != hipSuccess) return fail(h, DEM_ERR_HIP, "hipMalloc failed");
    tmp.push_back(*q);
    return DEM_OK;
  };
  dem::WallLocalArgs o{};
  o.ncol = nc;
  int *scan = nullptr;
  if ((rc = talloc((void **)&o.count, sizeof(int) * n)) || (rc = talloc((void **)&o.offset, sizeof(int) * (n + 1))) ||
      (rc = talloc((void **)&scan, sizeof(int) * dem::scan_tmp_size(n)))) {
    tfree();
    return rc;
  }
  dem::StepArgs a = step_args(h, h->cur_st, h->tri_cur);
  a.shearupdate = 0;
  hipError_t e = dem::launch_wall_local_count(a, o, h->stream);
  if (e == hipSuccess) e = dem::launch_exclusive_scan(o.count, o.offset, scan, n, h->stream);
  int total = 0;
  if (e == hipSuccess) e = hipMemcpyAsync(&total, o.offset + n, sizeof(int), hipMemcpyDeviceToHost, h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  if (e != hipSuccess) {
    tfree();
    return fail(h, DEM_ERR_HIP, hipGetErrorString(e));
  }
  *nrows = total;
  if (cap_out <= 0 || total == 0) {
    tfree();
    return DEM_OK;
  }
  if ((rc = talloc((void **)&o.ids, sizeof(int) * 3 * (size_t)total)) ||
      (rc = talloc((void **)&o.rec, sizeof(double) * (size_t)nc * (size_t)total))) {
    tfree();
    return rc;
  }
  e = dem::launch_wall_local_fill(a, o, h->stream);
  const int64_t m = std::min<int64_t>(cap_out, total);
  if (e == hipSuccess) e = hipMemcpyAsync(ids, o.ids, sizeof(int) * 3 * (size_t)m, hipMemcpyDeviceToHost, h->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(data, o.rec, sizeof(double) * (size_t)nc * (size_t)m, hipMemcpyDeviceToHost, h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  tfree();
  if (e != hipSuccess) return fail(h, DEM_ERR_HIP, hipGetErrorString(e));
  return DEM_OK;
}

int dem_get_pair_rolling(dem_handle *h, int64_t cap_out, double *rhist, int64_t *npairs) {
  if (!h || !npairs) return DEM_ERR_ARG;
  if (h->hw != 6) return fail(h, DEM_ERR_STATE, "no rolling history: no contact model uses epsd2");
  std::vector<int> tags(2 * std::max<int64_t>(cap_out, 1)), flags(std::max<int64_t>(cap_out, 1));
  return get_pairs(h, cap_out, tags.data(), flags.data(), rhist, npairs, 3);
}

int dem_get_wall_rolling(dem_handle *h, double *out) {
  if (!h) return DEM_ERR_ARG;
  if (h->hw != 6) return fail(h, DEM_ERR_STATE, "no rolling history: no contact model uses epsd2");
  return get_wall_history(h, out, 3);
}

int dem_get_stats(dem_handle *h, dem_stats *out) {
  if (!h || !out) return DEM_ERR_ARG;
  out->ntimestep = h->ntimestep;
  out->nrebuilds = h->nrebuilds;
  out->npairs = h->npairs;
  out->ncontacts = h->ncontacts;
  out->nwall_contacts = h->nwall_contacts;
  out->max_neighbors_seen = h->maxcount;
  out->nmesh_candidates = h->nmesh_cand;
  out->step_kernel_ms = h->kernel_launches ? h->kernel_ms / (double)h->kernel_launches : 0.0;
  out->step_kernel_launches = h->kernel_launches;
  out->rebuild_ms_total = h->rebuild_ms;
  return DEM_OK;
}

}  // extern "C"
