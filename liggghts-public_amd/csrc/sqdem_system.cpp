// Host side of the superquadric DEM step: the dem_sqdem_* C ABI. Owns the
// device arrays of one simulation, drives Verlet::setup / Verlet::run
// (verlet.cpp:134-205, 264-392) with the kernels of sqdem_kernels.hip, and
// optionally warm-starts new contacts from a trained network of a
// dem_sq_handle.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/liggghts_amd.h"
#include "dem_device.h"
#include "sq_kernels.h"
#include "sqdem_kernels.h"

namespace sqk {
const NetDev *handle_net(const dem_sq_handle *h, int *device);
}

namespace {

constexpr double kMyPi = 3.14159265358979323846;

struct List {
  int cap = 0;
  int np = 0;
  bool valid = false;
  int *first = nullptr, *pi = nullptr, *pj = nullptr, *flag = nullptr;
  double *hist = nullptr;
  std::vector<void *> allocs;  // capacity-dependent arrays
};

}  // namespace

struct dem_sqdem_handle {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev[2] = {nullptr, nullptr};
  std::string err;
  dem_sqdem_params p{};
  std::vector<void *> allocs;  // particle-count arrays
  // materials
  int ntypes = 0;
  bool mat_set = false;
  std::vector<double> hY, hG, hB, hmu;
  double *mat = nullptr;  // Y, G, B, mu tables, [ntypes+1]^2 each
  // particles
  int64_t n = 0;
  double *x = nullptr, *v = nullptr, *w = nullptr, *L = nullptr, *q = nullptr, *shape = nullptr, *blk = nullptr,
         *mass = nullptr, *rv = nullptr, *rb = nullptr, *I = nullptr, *f = nullptr, *t = nullptr, *xhold = nullptr,
         *wrap = nullptr, *whist = nullptr;
  int *type = nullptr;
  int maxtype = 0;
  // grid
  sqdk::Grid grid{};
  // lists
  List lists[2];
  int cur = 0;
  int *count = nullptr, *scan_tmp = nullptr, *rcount = nullptr, *rfirst = nullptr, *rfill = nullptr;
  int *overflow = nullptr, *dflag = nullptr;
  int pair_cap = 0;
  int *rlist = nullptr, *fresh_idx = nullptr;
  double *rec = nullptr;
  std::vector<void *> pair_allocs;
  unsigned long long *stats = nullptr, *totals = nullptr;
  int *hflag = nullptr;  // pinned
  // warm start
  dem_sq_handle *net = nullptr;
  int *fresh_count = nullptr;
  int64_t fresh_cap = 0;
  double *fpa = nullptr, *fpb = nullptr, *finit = nullptr;
  float *fX = nullptr, *fY = nullptr;
  std::vector<void *> fresh_allocs;
  // run state
  bool setup_done = false;
  int64_t steps = 0, rebuilds = 0;
  double run_ms = 0.0;
};

namespace {

std::string g_create_error;

int fail(dem_sqdem_handle *h, int code, const std::string &msg) {
  if (h)
    h->err = msg;
  else
    g_create_error = msg;
  return code;
}

#define SCHK(h, call)                                                             \
  do {                                                                            \
    hipError_t e_ = (call);                                                       \
    if (e_ != hipSuccess) return fail(h, DEM_ERR_HIP, hipGetErrorString(e_));     \
  } while (0)

int dalloc(dem_sqdem_handle *h, std::vector<void *> &list, void **p, size_t bytes) {
  *p = nullptr;
  if (bytes == 0) bytes = 16;
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) return fail(h, DEM_ERR_HIP, std::string("hipMalloc: ") + hipGetErrorString(e));
  list.push_back(*p);
  return DEM_OK;
}

void free_list(std::vector<void *> &list) {
  for (void *p : list) hipFree(p);
  list.clear();
}

// pow_abs_int / pow_abs / compDouble as in math_extra_liggghts_nonspherical.h:
// 137-221 and math_extra_liggghts.h:536-547
double pow_abs_int(double a, int b) {
  a = std::fabs(a);
  double r = 1.0;
  switch (b) {
    case 0: return 1.0;
    case 1: return a;
    case 2: return a * a;
    case 3: return a * a * a;
    case 4: { const double a2 = a * a; return a2 * a2; }
    case 5: { const double a2 = a * a; return a2 * a2 * a; }
    case 6: { const double a3 = a * a * a; return a3 * a3; }
    case 7: { const double a2 = a * a, a3 = a2 * a; return a2 * a2 * a3; }
    case 8: { const double a2 = a * a, a4 = a2 * a2; return a4 * a4; }
    case 9: { const double a3 = a * a * a; return a3 * a3 * a3; }
    default:
      r = a;
      for (int i = 0; i < b - 1; i++) r *= a;
      return r;
  }
}

double pow_abs(double base, double exponent) {
  base = std::fabs(base);
  const double tol = 1e-12;
  if (std::fabs(exponent) < tol || std::fabs(base - 1.0) < tol) return 1.0;
  if (std::fabs(base) < tol) return 0.0;
  const double n = std::floor(std::fabs(exponent));
  double r = pow_abs_int(base, (int)n);
  const double nd = std::fabs(exponent) - n;
  if (nd > 0.0) {
    if (std::fabs(2.0 * nd - 1.0) < 2.0 * tol)
      r *= std::sqrt(base);
    else
      r *= std::pow(base, nd);
  }
  return exponent > 0.0 ? r : 1.0 / r;
}

bool comp_double(double a, double b, double prec) {
  if (a == b) return true;
  if (b == 0) return a < prec && a > -prec;
  const double x = a - b;
  return x < prec && x > -prec;
}

// bounding_sphere_radius_superquadric (math_extra_liggghts_superquadric.cpp:206-236)
double bounding_radius(const double *shape, const double *blk) {
  double r;
  if (comp_double(blk[0], 2.0, 1e-13) && comp_double(blk[1], 2.0, 1e-13)) {
    r = std::max(std::max(shape[0], shape[1]), shape[2]);
  } else {
    const double a = std::max(shape[0], shape[1]);
    const double b = std::min(shape[0], shape[1]);
    const double c = shape[2];
    if (comp_double(blk[1], 2.0, 1e-13)) {
      r = std::sqrt(a * a + c * c);
    } else {
      const double n1 = std::max(2.1, blk[0]);
      const double n2 = std::max(2.1, blk[1]);
      const double alpha = pow_abs(b / a, 2.0 / (n2 - 2.0));
      const double gamma = pow_abs(1.0 + pow_abs(alpha, n2), n1 / n2 - 1.0);
      const double beta = pow_abs(c * c / (a * a) * gamma, 1.0 / (n1 - 2.0));
      const double x0 = 1.0 / pow_abs(pow_abs(1.0 + pow_abs(alpha, n2), n1 / n2) + pow_abs(beta, n1), 1.0 / n1);
      const double xt = x0 * a, yt = alpha * x0 * b, zt = beta * x0 * c;
      r = std::sqrt(xt * xt + yt * yt + zt * zt);
    }
  }
  return r * 1.0001;  // scale_koef, math_extra_liggghts_nonspherical.h:60
}

double beta_fn(double a, double b) { return std::exp(std::lgamma(a) + std::lgamma(b) - std::lgamma(a + b)); }

// volume_superquadric / inertia_superquadric (math_extra_liggghts_superquadric.cpp:242-268)
void mass_props(const double *shape, const double *blk, double dens, double *volume, double *I) {
  const double eps1 = 2.0 / blk[0], eps2 = 2.0 / blk[1];
  const double a1 = shape[0], a2 = shape[1], a3 = shape[2];
  *volume = 2.0 * a1 * a2 * a3 * eps1 * eps2 * beta_fn(0.5 * eps1, eps1 + 1.0) * beta_fn(0.5 * eps2, 0.5 * eps2 + 1.0);
  const double bA = beta_fn(1.5 * eps2, 0.5 * eps2) * beta_fn(0.5 * eps1, 2.0 * eps1 + 1.0);
  const double bB = beta_fn(0.5 * eps2, 0.5 * eps2 + 1.0) * beta_fn(1.5 * eps1, eps1 + 1.0);
  I[0] = 0.5 * a1 * a2 * a3 * eps1 * eps2 * (a2 * a2 * bA + 4.0 * a3 * a3 * bB) * dens;
  I[1] = 0.5 * a1 * a2 * a3 * eps1 * eps2 * (a1 * a1 * bA + 4.0 * a3 * a3 * bB) * dens;
  I[2] = 0.5 * a1 * a2 * a3 * eps1 * eps2 * (a1 * a1 + a2 * a2) * bA * dens;
}

void rot_of(const double *q, double *m) {
  const double w2 = q[0] * q[0], i2 = q[1] * q[1], j2 = q[2] * q[2], k2 = q[3] * q[3];
  const double twoij = 2.0 * q[1] * q[2], twoik = 2.0 * q[1] * q[3], twojk = 2.0 * q[2] * q[3];
  const double twoiw = 2.0 * q[1] * q[0], twojw = 2.0 * q[2] * q[0], twokw = 2.0 * q[3] * q[0];
  m[0] = w2 + i2 - j2 - k2;
  m[1] = twoij - twokw;
  m[2] = twojw + twoik;
  m[3] = twoij + twokw;
  m[4] = w2 - i2 + j2 - k2;
  m[5] = twojk - twoiw;
  m[6] = twoik - twojw;
  m[7] = twojk + twoiw;
  m[8] = w2 - i2 - j2 + k2;
}

sqdk::Dev make_dev(dem_sqdem_handle *h) {
  sqdk::Dev d{};
  d.n = (int)h->n;
  d.x = h->x;
  d.v = h->v;
  d.w = h->w;
  d.L = h->L;
  d.q = h->q;
  d.shape = h->shape;
  d.blk = h->blk;
  d.mass = h->mass;
  d.rv = h->rv;
  d.rb = h->rb;
  d.I = h->I;
  d.f = h->f;
  d.t = h->t;
  d.xhold = h->xhold;
  d.wrap = h->wrap;
  d.type = h->type;
  for (int k = 0; k < 3; k++) {
    d.lo[k] = h->p.boxlo[k];
    d.hi[k] = h->p.boxhi[k];
    d.len[k] = h->p.boxhi[k] - h->p.boxlo[k];
    d.per[k] = h->p.periodic[k];
  }
  d.dt = h->p.dt;
  d.skin = h->p.skin;
  d.tdamp = h->p.tangential_damping;
  d.limit = h->p.limit_force;
  d.t1 = h->ntypes + 1;
  const size_t tt = (size_t)d.t1 * d.t1;
  d.Y = h->mat;
  d.G = h->mat + tt;
  d.B = h->mat + 2 * tt;
  d.mu = h->mat + 3 * tt;
  d.nplanes = h->p.nplanes;
  for (int k = 0; k < h->p.nplanes; k++) {
    d.pax[k] = h->p.plane_axis[k];
    d.ppos[k] = h->p.plane_pos[k];
    d.ptype[k] = h->p.plane_type[k];
  }
  d.has_grav = h->p.has_gravity;
  d.grav_first = h->p.gravity_first;
  for (int k = 0; k < 3; k++) d.g[k] = h->p.gravity[k];
  const List &l = h->lists[h->cur];
  d.np = l.valid ? l.np : 0;
  d.first = l.first;
  d.pi = l.pi;
  d.pj = l.pj;
  d.flag = l.flag;
  d.hist = l.hist;
  d.rec = h->rec;
  d.rfirst = h->rfirst;
  d.rlist = h->rlist;
  d.whist = h->whist;
  d.stats = h->stats;
  d.init = nullptr;
  d.fresh_idx = h->fresh_idx;
  d.max_iter = 100000;  // Niter of calc_contact_point (math_extra_liggghts_superquadric.cpp:856)
  return d;
}

int ensure_list(dem_sqdem_handle *h, List &l, int np) {
  if (np <= l.cap) return DEM_OK;
  SCHK(h, hipStreamSynchronize(h->stream));
  const int cap = std::max(np + np / 4 + 1024, 4096);
  // first[] survives: it is sized by the particle count
  std::vector<void *> keep;
  for (void *p : l.allocs)
    if (p == l.first) keep.push_back(p);
    else hipFree(p);
  l.allocs = keep;
  int rc;
  if ((rc = dalloc(h, l.allocs, (void **)&l.pi, sizeof(int) * cap)) ||
      (rc = dalloc(h, l.allocs, (void **)&l.pj, sizeof(int) * cap)) ||
      (rc = dalloc(h, l.allocs, (void **)&l.flag, sizeof(int) * cap)) ||
      (rc = dalloc(h, l.allocs, (void **)&l.hist, sizeof(double) * sqdk::kHist * (size_t)cap)))
    return rc;
  l.cap = cap;
  return DEM_OK;
}

int ensure_pair_arrays(dem_sqdem_handle *h, int np) {
  if (np <= h->pair_cap) return DEM_OK;
  SCHK(h, hipStreamSynchronize(h->stream));
  free_list(h->pair_allocs);
  const int cap = std::max(np + np / 4 + 1024, 4096);
  int rc;
  if ((rc = dalloc(h, h->pair_allocs, (void **)&h->rlist, sizeof(int) * cap)) ||
      (rc = dalloc(h, h->pair_allocs, (void **)&h->fresh_idx, sizeof(int) * cap)) ||
      (rc = dalloc(h, h->pair_allocs, (void **)&h->rec, sizeof(double) * sqdk::kRec * (size_t)cap)))
    return rc;
  h->pair_cap = cap;
  return DEM_OK;
}

// Neighbor::build with the history carried over (neigh_gran.cpp:485-639,
// fix_contact_history.cpp:305-425), after Domain::pbc
int rebuild(dem_sqdem_handle *h) {
  const int old = h->cur;
  h->cur ^= 1;
  List &nl = h->lists[h->cur];
  const List &ol = h->lists[old];
  sqdk::Dev d = make_dev(h);
  SCHK(h, sqdk::launch_wrap(d, h->stream));
  SCHK(h, sqdk::launch_bin(d, h->grid, h->stream));
  sqdk::Build b{};
  b.count = h->count;
  SCHK(h, sqdk::launch_pair_count(d, h->grid, b, h->stream));
  SCHK(h, dem::launch_exclusive_scan(h->count, nl.first, h->scan_tmp, (int)h->n, h->stream));
  int np = 0;
  SCHK(h, hipMemcpyAsync(h->hflag, nl.first + h->n, sizeof(int), hipMemcpyDeviceToHost, h->stream));
  SCHK(h, hipStreamSynchronize(h->stream));
  np = *h->hflag;
  if (np < 0) return fail(h, DEM_ERR_OVERFLOW, "pair list exceeds 2^31 entries");
  int rc;
  if ((rc = ensure_list(h, nl, np)) || (rc = ensure_pair_arrays(h, np))) return rc;
  b.first = nl.first;
  b.pi = nl.pi;
  b.pj = nl.pj;
  b.flag = nl.flag;
  b.hist = nl.hist;
  b.cap = nl.cap;
  if (ol.valid) {
    b.ofirst = ol.first;
    b.opj = ol.pj;
    b.oflag = ol.flag;
    b.ohist = ol.hist;
    b.onp = ol.np;
  }
  b.rcount = h->rcount;
  b.rfirst = h->rfirst;
  b.rfill = h->rfill;
  b.rlist = h->rlist;
  b.overflow = h->overflow;
  SCHK(h, hipMemsetAsync(h->overflow, 0, sizeof(int), h->stream));
  SCHK(h, sqdk::launch_pair_fill(d, h->grid, b, h->stream));
  nl.np = np;
  nl.valid = true;
  d = make_dev(h);
  SCHK(h, sqdk::launch_reverse(d, b, h->scan_tmp, h->stream));
  SCHK(h, hipMemcpyAsync(h->hflag, h->overflow, sizeof(int), hipMemcpyDeviceToHost, h->stream));
  SCHK(h, hipStreamSynchronize(h->stream));
  if (*h->hflag) return fail(h, DEM_ERR_OVERFLOW, "pair list overflow");
  h->rebuilds++;
  return DEM_OK;
}

int ensure_fresh(dem_sqdem_handle *h, int64_t m) {
  if (m <= h->fresh_cap) return DEM_OK;
  SCHK(h, hipStreamSynchronize(h->stream));
  free_list(h->fresh_allocs);
  const int64_t cap = std::max<int64_t>(m + m / 4 + 256, 1024);
  int rc;
  if ((rc = dalloc(h, h->fresh_allocs, (void **)&h->fpa, sizeof(double) * 12 * (size_t)cap)) ||
      (rc = dalloc(h, h->fresh_allocs, (void **)&h->fpb, sizeof(double) * 12 * (size_t)cap)) ||
      (rc = dalloc(h, h->fresh_allocs, (void **)&h->finit, sizeof(double) * 3 * (size_t)cap)) ||
      (rc = dalloc(h, h->fresh_allocs, (void **)&h->fX, sizeof(float) * sqk::kIn * (size_t)cap)) ||
      (rc = dalloc(h, h->fresh_allocs, (void **)&h->fY, sizeof(float) * sqk::kOut * (size_t)cap)))
    return rc;
  h->fresh_cap = cap;
  return DEM_OK;
}

// force_clear, pair, post_force, and (integrate) final_integrate
int forces(dem_sqdem_handle *h, int shearupdate, int integrate) {
  sqdk::Dev d = make_dev(h);
  SCHK(h, sqdk::launch_zero_stats(h->stats, h->stream));
  if (h->net && d.np > 0) {
    const sqk::NetDev *net = sqk::handle_net(h->net, nullptr);
    if (!net) return fail(h, DEM_ERR_STATE, "warm-start network not initialised");
    int rc;
    // the compact pose arrays hold at most one row per pair
    if ((rc = ensure_fresh(h, d.np))) return rc;
    SCHK(h, sqdk::launch_fresh_scan(d, h->fresh_idx, h->fresh_count, h->fpa, h->fpb, h->stream));
    SCHK(h, hipMemcpyAsync(h->hflag, h->fresh_count, sizeof(int), hipMemcpyDeviceToHost, h->stream));
    SCHK(h, hipStreamSynchronize(h->stream));
    const int m = *h->hflag;
    if (m > 0) {
      SCHK(h, sqk::launch_features(h->fpa, h->fpb, m, h->fX, h->stream));
      SCHK(h, sqk::launch_mlp_forward(*net, h->fX, m, h->fY, h->stream));
      SCHK(h, sqk::launch_to_point(h->fpa, h->fpb, m, h->fY, h->finit, h->stream));
    }
    d.init = h->finit;
  }
  SCHK(h, sqdk::launch_pairs(d, shearupdate, h->stream));
  SCHK(h, sqdk::launch_gather(d, shearupdate, integrate, h->stream));
  SCHK(h, sqdk::launch_accumulate(h->stats, h->totals, h->stream));
  return DEM_OK;
}

int setup(dem_sqdem_handle *h) {
  int rc;
  if ((rc = rebuild(h))) return rc;
  if ((rc = forces(h, 0, 0))) return rc;
  h->setup_done = true;
  return DEM_OK;
}

void free_particles(dem_sqdem_handle *h) {
  free_list(h->allocs);
  for (auto &l : h->lists) {
    free_list(l.allocs);
    l = List();
  }
  free_list(h->pair_allocs);
  h->pair_cap = 0;
  free_list(h->fresh_allocs);
  h->fresh_cap = 0;
  h->n = 0;
  h->mat = nullptr;
}

int upload_materials(dem_sqdem_handle *h) {
  const size_t tt = (size_t)(h->ntypes + 1) * (h->ntypes + 1);
  std::vector<double> blob(4 * tt);
  std::copy(h->hY.begin(), h->hY.end(), blob.begin());
  std::copy(h->hG.begin(), h->hG.end(), blob.begin() + tt);
  std::copy(h->hB.begin(), h->hB.end(), blob.begin() + 2 * tt);
  std::copy(h->hmu.begin(), h->hmu.end(), blob.begin() + 3 * tt);
  int rc;
  if ((rc = dalloc(h, h->allocs, (void **)&h->mat, sizeof(double) * blob.size()))) return rc;
  SCHK(h, hipMemcpy(h->mat, blob.data(), sizeof(double) * blob.size(), hipMemcpyHostToDevice));
  return DEM_OK;
}

}  // namespace

extern "C" {

int dem_sqdem_create(int device, const dem_sqdem_params *p, dem_sqdem_handle **out) {
  if (!out || !p) return fail(nullptr, DEM_ERR_ARG, "null argument");
  *out = nullptr;
  if (!(p->dt > 0.0) || !std::isfinite(p->dt)) return fail(nullptr, DEM_ERR_ARG, "timestep must be > 0");
  if (!(p->skin >= 0.0) || !std::isfinite(p->skin)) return fail(nullptr, DEM_ERR_ARG, "skin must be >= 0");
  for (int k = 0; k < 3; k++)
    if (!(p->boxhi[k] > p->boxlo[k])) return fail(nullptr, DEM_ERR_ARG, "empty box");
  if (p->nplanes < 0 || p->nplanes > sqdk::kMaxPlanes) return fail(nullptr, DEM_ERR_ARG, "at most 6 planes");
  for (int k = 0; k < p->nplanes; k++) {
    if (p->plane_axis[k] < 0 || p->plane_axis[k] > 2) return fail(nullptr, DEM_ERR_ARG, "plane axis is 0, 1 or 2");
    if (p->plane_type[k] < 1) return fail(nullptr, DEM_ERR_ARG, "wall type must be >= 1");
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev)
    return fail(nullptr, DEM_ERR_NODEV, "no HIP device available");
  if (hipSetDevice(device) != hipSuccess) return fail(nullptr, DEM_ERR_NODEV, "hipSetDevice failed");
  dem_sqdem_handle *h = new dem_sqdem_handle();
  h->device = device;
  h->p = *p;
  auto bad = [&](const char *what) {
    g_create_error = what;
    dem_sqdem_destroy(h);
    return DEM_ERR_HIP;
  };
  if (hipStreamCreate(&h->stream) != hipSuccess) return bad("hipStreamCreate failed");
  for (auto &e : h->ev)
    if (hipEventCreate(&e) != hipSuccess) return bad("hipEventCreate failed");
  if (hipHostMalloc((void **)&h->hflag, sizeof(int) * 4, 0) != hipSuccess) return bad("hipHostMalloc failed");
  std::vector<void *> &fixed = h->fresh_allocs;  // freed with the handle; re-created below when reset
  (void)fixed;
  if (hipMalloc((void **)&h->stats, sizeof(unsigned long long) * sqdk::kNumStats * 2) != hipSuccess)
    return bad("hipMalloc failed");
  h->totals = h->stats + sqdk::kNumStats;
  if (hipMemset(h->stats, 0, sizeof(unsigned long long) * sqdk::kNumStats * 2) != hipSuccess)
    return bad("hipMemset failed");
  if (hipMalloc((void **)&h->fresh_count, sizeof(int) * 4) != hipSuccess) return bad("hipMalloc failed");
  *out = h;
  return DEM_OK;
}

int dem_sqdem_destroy(dem_sqdem_handle *h) {
  if (!h) return DEM_OK;
  if (h->stream) hipStreamSynchronize(h->stream);
  free_particles(h);
  if (h->stats) hipFree(h->stats);
  if (h->fresh_count) hipFree(h->fresh_count);
  if (h->hflag) hipHostFree(h->hflag);
  for (auto &e : h->ev)
    if (e) hipEventDestroy(e);
  if (h->stream) hipStreamDestroy(h->stream);
  delete h;
  return DEM_OK;
}

const char *dem_sqdem_last_error(const dem_sqdem_handle *h) { return h ? h->err.c_str() : g_create_error.c_str(); }

int dem_sqdem_set_materials(dem_sqdem_handle *h, int ntypes, const double *Y, const double *nu, const double *e,
                            const double *mu) {
  if (!h) return DEM_ERR_ARG;
  if (ntypes < 1 || ntypes > 64 || !Y || !nu || !e || !mu) return fail(h, DEM_ERR_ARG, "bad material arguments");
  // range checks of global_properties.cpp:316-320, 360-372, 393-396
  for (int i = 0; i < ntypes; i++) {
    if (Y[i] < 5e6) return fail(h, DEM_ERR_ARG, "youngsModulus >= 5e6 required for SI units");
    if (Y[i] > 1e9) return fail(h, DEM_ERR_ARG, "youngsModulus <= 1e9 required for SI units");
    if (nu[i] < 0. || nu[i] > 0.5) return fail(h, DEM_ERR_ARG, "0 <= poissonsRatio <= 0.5 required");
  }
  for (int k = 0; k < ntypes * ntypes; k++)
    if (e[k] <= 0.05 || e[k] > 1) return fail(h, DEM_ERR_ARG, "0.05 < coefficientRestitution <= 1 required");
  for (int k = 0; k < h->p.nplanes; k++)
    if (h->p.plane_type[k] > ntypes) return fail(h, DEM_ERR_ARG, "wall type beyond the atom types");
  if (h->n > 0 && h->maxtype > ntypes) return fail(h, DEM_ERR_STATE, "fewer atom types than the particles use");
  const int t1 = ntypes + 1;
  h->ntypes = ntypes;
  h->hY.assign((size_t)t1 * t1, 0.0);
  h->hG.assign((size_t)t1 * t1, 0.0);
  h->hB.assign((size_t)t1 * t1, 0.0);
  h->hmu.assign((size_t)t1 * t1, 0.0);
  // global_properties.cpp:426-433, 455-464, 518-541
  for (int i = 1; i <= ntypes; i++)
    for (int j = 1; j <= ntypes; j++) {
      const double Yi = Y[i - 1], Yj = Y[j - 1], vi = nu[i - 1], vj = nu[j - 1];
      const size_t k = (size_t)i * t1 + j;
      h->hY[k] = 1. / ((1. - std::pow(vi, 2.)) / Yi + (1. - std::pow(vj, 2.)) / Yj);
      h->hG[k] = 1. / (2. * (2. - vi) * (1. + vi) / Yi + 2. * (2. - vj) * (1. + vj) / Yj);
      const double le = std::log(e[(i - 1) * ntypes + (j - 1)]);
      h->hB[k] = le / std::sqrt(std::pow(le, 2.) + std::pow(kMyPi, 2.));
      h->hmu[k] = mu[(i - 1) * ntypes + (j - 1)];
    }
  h->mat_set = true;
  if (h->n > 0) {
    SCHK(h, hipSetDevice(h->device));
    SCHK(h, hipStreamSynchronize(h->stream));
    return upload_materials(h);
  }
  return DEM_OK;
}

int dem_sqdem_set_particles(dem_sqdem_handle *h, int64_t n, const double *x, const double *v, const double *omega,
                            const double *quat, const double *shape, const double *blockiness,
                            const double *density, const int *type) {
  if (!h) return DEM_ERR_ARG;
  if (!h->mat_set) return fail(h, DEM_ERR_STATE, "set materials first");
  if (n <= 0 || n > (1 << 28) || !x || !v || !omega || !quat || !shape || !blockiness || !density || !type)
    return fail(h, DEM_ERR_ARG, "bad particle arrays");
  const size_t N = (size_t)n;
  std::vector<double> mass(N), rv(N), rb(N), I(3 * N), L(3 * N);
  int maxtype = 0;
  double rbmax = 0.0;
  for (size_t i = 0; i < N; i++) {
    for (int k = 0; k < 3; k++)
      if (!std::isfinite(x[3 * i + k]) || !std::isfinite(v[3 * i + k]) || !std::isfinite(omega[3 * i + k]))
        return fail(h, DEM_ERR_ARG, "non-finite particle state");
    // AtomVecSuperquadric::data_atom checks (atom_vec_superquadric.cpp:1163-1184)
    if (!(shape[3 * i] > 0 && shape[3 * i + 1] > 0 && shape[3 * i + 2] > 0))
      return fail(h, DEM_ERR_ARG, "Invalid shape in Atoms section of data file");
    if (!(blockiness[2 * i] >= 2.0 && blockiness[2 * i + 1] >= 2.0))
      return fail(h, DEM_ERR_ARG, "Invalid blockiness in Atoms section of data file");
    if (!(density[i] > 0.0)) return fail(h, DEM_ERR_ARG, "Invalid density in Atoms section of data file");
    if (type[i] < 1 || type[i] > h->ntypes) return fail(h, DEM_ERR_ARG, "Invalid atom type in Atoms section of data file");
    const double *qq = quat + 4 * i;
    const double qn = qq[0] * qq[0] + qq[1] * qq[1] + qq[2] * qq[2] + qq[3] * qq[3];
    if (std::fabs(qn - 1.0) > 1e-6) return fail(h, DEM_ERR_ARG, "quaternions must be unit");
    maxtype = std::max(maxtype, type[i]);
    double vol;
    mass_props(shape + 3 * i, blockiness + 2 * i, density[i], &vol, &I[3 * i]);
    mass[i] = vol * density[i];
    rv[i] = std::cbrt(0.75 * vol / kMyPi);
    rb[i] = bounding_radius(shape + 3 * i, blockiness + 2 * i);
    rbmax = std::max(rbmax, rb[i]);
    // angular momentum of the given angular velocity, R diag(I) R^T omega
    double R[9], wb[3], mb[3];
    rot_of(qq, R);
    const double *wi = omega + 3 * i;
    wb[0] = R[0] * wi[0] + R[3] * wi[1] + R[6] * wi[2];
    wb[1] = R[1] * wi[0] + R[4] * wi[1] + R[7] * wi[2];
    wb[2] = R[2] * wi[0] + R[5] * wi[1] + R[8] * wi[2];
    for (int k = 0; k < 3; k++) mb[k] = I[3 * i + k] * wb[k];
    for (int k = 0; k < 3; k++) L[3 * i + k] = R[3 * k] * mb[0] + R[3 * k + 1] * mb[1] + R[3 * k + 2] * mb[2];
  }
  // grid: cells no smaller than the largest pair cutoff 2 rb_max + skin
  const double cut = 2.0 * rbmax + h->p.skin;
  sqdk::Grid g{};
  int64_t ncell = 1;
  for (int k = 0; k < 3; k++) {
    const double len = h->p.boxhi[k] - h->p.boxlo[k];
    int nc = (int)std::floor(len / cut);
    if (nc < 1) nc = 1;
    if (h->p.periodic[k] && nc < 3)
      return fail(h, DEM_ERR_ARG, "periodic box shorter than three neighbor cutoffs");
    if (nc > 4096) nc = 4096;
    g.nc[k] = nc;
    g.cs[k] = len / nc;
    ncell *= nc;
  }
  if (ncell > (1 << 26)) return fail(h, DEM_ERR_ARG, "too many neighbor cells");
  g.ncell = (int)ncell;
  SCHK(h, hipSetDevice(h->device));
  SCHK(h, hipStreamSynchronize(h->stream));
  free_particles(h);
  h->n = n;
  h->maxtype = maxtype;
  h->setup_done = false;
  h->steps = h->rebuilds = 0;
  SCHK(h, hipMemset(h->stats, 0, sizeof(unsigned long long) * sqdk::kNumStats * 2));
  int rc;
  auto A = [&](void **p, size_t b) { return dalloc(h, h->allocs, p, b); };
  const size_t D3 = sizeof(double) * 3 * N;
  const int nplanes = std::max(h->p.nplanes, 1);
  if ((rc = A((void **)&h->x, D3)) || (rc = A((void **)&h->v, D3)) || (rc = A((void **)&h->w, D3)) ||
      (rc = A((void **)&h->L, D3)) || (rc = A((void **)&h->q, sizeof(double) * 4 * N)) ||
      (rc = A((void **)&h->shape, D3)) || (rc = A((void **)&h->blk, sizeof(double) * 2 * N)) ||
      (rc = A((void **)&h->mass, sizeof(double) * N)) || (rc = A((void **)&h->rv, sizeof(double) * N)) ||
      (rc = A((void **)&h->rb, sizeof(double) * N)) || (rc = A((void **)&h->I, D3)) || (rc = A((void **)&h->f, D3)) ||
      (rc = A((void **)&h->t, D3)) || (rc = A((void **)&h->xhold, D3)) || (rc = A((void **)&h->wrap, D3)) ||
      (rc = A((void **)&h->whist, sizeof(double) * 3 * N * nplanes)) || (rc = A((void **)&h->type, sizeof(int) * N)) ||
      (rc = A((void **)&g.cell_of, sizeof(int) * N)) || (rc = A((void **)&g.count, sizeof(int) * (ncell + 1))) ||
      (rc = A((void **)&g.start, sizeof(int) * (ncell + 1))) || (rc = A((void **)&g.fill, sizeof(int) * (ncell + 1))) ||
      (rc = A((void **)&g.list, sizeof(int) * N)) ||
      (rc = A((void **)&h->scan_tmp, sizeof(int) * dem::scan_tmp_size((int)std::max<int64_t>(ncell, n + 1)))) ||
      (rc = A((void **)&h->count, sizeof(int) * (N + 1))) || (rc = A((void **)&h->rcount, sizeof(int) * (N + 1))) ||
      (rc = A((void **)&h->rfirst, sizeof(int) * (N + 1))) || (rc = A((void **)&h->rfill, sizeof(int) * (N + 1))) ||
      (rc = A((void **)&h->overflow, sizeof(int) * 4)) || (rc = A((void **)&h->dflag, sizeof(int) * 4)))
    return rc;
  g.scan_tmp = h->scan_tmp;
  h->grid = g;
  for (auto &l : h->lists)
    if ((rc = dalloc(h, l.allocs, (void **)&l.first, sizeof(int) * (N + 1)))) return rc;
  SCHK(h, hipMemcpy(h->x, x, D3, hipMemcpyHostToDevice));
  SCHK(h, hipMemcpy(h->v, v, D3, hipMemcpyHostToDevice));
  SCHK(h, hipMemcpy(h->w, omega, D3, hipMemcpyHostToDevice));
  SCHK(h, hipMemcpy(h->L, L.data(), D3, hipMemcpyHostToDevice));
  SCHK(h, hipMemcpy(h->q, quat, sizeof(double) * 4 * N, hipMemcpyHostToDevice));
  SCHK(h, hipMemcpy(h->shape, shape, D3, hipMemcpyHostToDevice));
  SCHK(h, hipMemcpy(h->blk, blockiness, sizeof(double) * 2 * N, hipMemcpyHostToDevice));
  SCHK(h, hipMemcpy(h->mass, mass.data(), sizeof(double) * N, hipMemcpyHostToDevice));
  SCHK(h, hipMemcpy(h->rv, rv.data(), sizeof(double) * N, hipMemcpyHostToDevice));
  SCHK(h, hipMemcpy(h->rb, rb.data(), sizeof(double) * N, hipMemcpyHostToDevice));
  SCHK(h, hipMemcpy(h->I, I.data(), D3, hipMemcpyHostToDevice));
  SCHK(h, hipMemcpy(h->type, type, sizeof(int) * N, hipMemcpyHostToDevice));
  SCHK(h, hipMemset(h->f, 0, D3));
  SCHK(h, hipMemset(h->t, 0, D3));
  SCHK(h, hipMemset(h->whist, 0, sizeof(double) * 3 * N * nplanes));
  return upload_materials(h);
}

int dem_sqdem_set_warm_start(dem_sqdem_handle *h, dem_sq_handle *net) {
  if (!h) return DEM_ERR_ARG;
  if (net) {
    int dev = -1;
    if (!sqk::handle_net(net, &dev)) return fail(h, DEM_ERR_STATE, "the solver handle has no network");
    if (dev != h->device) return fail(h, DEM_ERR_ARG, "network on another device");
  }
  h->net = net;
  return DEM_OK;
}

int dem_sqdem_setup(dem_sqdem_handle *h) {
  if (!h) return DEM_ERR_ARG;
  if (h->n <= 0) return fail(h, DEM_ERR_STATE, "no particles");
  SCHK(h, hipSetDevice(h->device));
  return setup(h);
}

int dem_sqdem_run(dem_sqdem_handle *h, int nsteps) {
  if (!h) return DEM_ERR_ARG;
  if (h->n <= 0) return fail(h, DEM_ERR_STATE, "no particles");
  if (nsteps < 0) return fail(h, DEM_ERR_ARG, "negative step count");
  SCHK(h, hipSetDevice(h->device));
  SCHK(h, hipEventRecord(h->ev[0], h->stream));
  int rc;
  if (!h->setup_done && (rc = setup(h))) return rc;
  for (int s = 0; s < nsteps; s++) {
    sqdk::Dev d = make_dev(h);
    SCHK(h, hipMemsetAsync(h->dflag, 0, sizeof(int), h->stream));
    SCHK(h, sqdk::launch_initial(d, h->dflag, h->stream));
    SCHK(h, hipMemcpyAsync(h->hflag, h->dflag, sizeof(int), hipMemcpyDeviceToHost, h->stream));
    SCHK(h, hipStreamSynchronize(h->stream));
    if (*h->hflag && (rc = rebuild(h))) return rc;
    if ((rc = forces(h, 1, 1))) return rc;
    h->steps++;
  }
  SCHK(h, hipEventRecord(h->ev[1], h->stream));
  SCHK(h, hipEventSynchronize(h->ev[1]));
  float ms = 0.f;
  SCHK(h, hipEventElapsedTime(&ms, h->ev[0], h->ev[1]));
  h->run_ms = ms;
  return DEM_OK;
}

int dem_sqdem_get_state(dem_sqdem_handle *h, double *x, double *v, double *omega, double *angmom, double *quat,
                        double *f, double *torque) {
  if (!h) return DEM_ERR_ARG;
  if (h->n <= 0) return fail(h, DEM_ERR_STATE, "no particles");
  SCHK(h, hipStreamSynchronize(h->stream));
  const size_t D3 = sizeof(double) * 3 * (size_t)h->n;
  if (x) SCHK(h, hipMemcpy(x, h->x, D3, hipMemcpyDeviceToHost));
  if (v) SCHK(h, hipMemcpy(v, h->v, D3, hipMemcpyDeviceToHost));
  if (omega) SCHK(h, hipMemcpy(omega, h->w, D3, hipMemcpyDeviceToHost));
  if (angmom) SCHK(h, hipMemcpy(angmom, h->L, D3, hipMemcpyDeviceToHost));
  if (quat) SCHK(h, hipMemcpy(quat, h->q, sizeof(double) * 4 * (size_t)h->n, hipMemcpyDeviceToHost));
  if (f) SCHK(h, hipMemcpy(f, h->f, D3, hipMemcpyDeviceToHost));
  if (torque) SCHK(h, hipMemcpy(torque, h->t, D3, hipMemcpyDeviceToHost));
  return DEM_OK;
}

int dem_sqdem_get_props(dem_sqdem_handle *h, double *mass, double *rvol, double *rbound, double *inertia) {
  if (!h) return DEM_ERR_ARG;
  if (h->n <= 0) return fail(h, DEM_ERR_STATE, "no particles");
  SCHK(h, hipStreamSynchronize(h->stream));
  const size_t N = (size_t)h->n;
  if (mass) SCHK(h, hipMemcpy(mass, h->mass, sizeof(double) * N, hipMemcpyDeviceToHost));
  if (rvol) SCHK(h, hipMemcpy(rvol, h->rv, sizeof(double) * N, hipMemcpyDeviceToHost));
  if (rbound) SCHK(h, hipMemcpy(rbound, h->rb, sizeof(double) * N, hipMemcpyDeviceToHost));
  if (inertia) SCHK(h, hipMemcpy(inertia, h->I, sizeof(double) * 3 * N, hipMemcpyDeviceToHost));
  return DEM_OK;
}

int64_t dem_sqdem_get_pairs(dem_sqdem_handle *h, int64_t cap, int *ij, int *flags, double *hist) {
  if (!h) return DEM_ERR_ARG;
  if (h->n <= 0) return fail(h, DEM_ERR_STATE, "no particles");
  SCHK(h, hipStreamSynchronize(h->stream));
  const List &l = h->lists[h->cur];
  const int64_t np = l.valid ? l.np : 0;
  if (cap < np) return -np;
  if (np == 0) return 0;
  if (ij) {
    std::vector<int> a(np), b(np);
    SCHK(h, hipMemcpy(a.data(), l.pi, sizeof(int) * np, hipMemcpyDeviceToHost));
    SCHK(h, hipMemcpy(b.data(), l.pj, sizeof(int) * np, hipMemcpyDeviceToHost));
    for (int64_t k = 0; k < np; k++) {
      ij[2 * k] = a[k];
      ij[2 * k + 1] = b[k];
    }
  }
  if (flags) SCHK(h, hipMemcpy(flags, l.flag, sizeof(int) * np, hipMemcpyDeviceToHost));
  if (hist) SCHK(h, hipMemcpy(hist, l.hist, sizeof(double) * sqdk::kHist * np, hipMemcpyDeviceToHost));
  return np;
}

int dem_sqdem_get_wall_history(dem_sqdem_handle *h, double *hist) {
  if (!h || !hist) return DEM_ERR_ARG;
  if (h->n <= 0) return fail(h, DEM_ERR_STATE, "no particles");
  SCHK(h, hipStreamSynchronize(h->stream));
  if (h->p.nplanes > 0)
    SCHK(h, hipMemcpy(hist, h->whist, sizeof(double) * 3 * (size_t)h->n * h->p.nplanes, hipMemcpyDeviceToHost));
  return DEM_OK;
}

int dem_sqdem_get_stats(dem_sqdem_handle *h, dem_sqdem_stats *out) {
  if (!h || !out) return DEM_ERR_ARG;
  SCHK(h, hipStreamSynchronize(h->stream));
  unsigned long long s[2 * sqdk::kNumStats];
  SCHK(h, hipMemcpy(s, h->stats, sizeof(s), hipMemcpyDeviceToHost));
  const unsigned long long *tot = s + sqdk::kNumStats;
  std::memset(out, 0, sizeof(*out));
  out->steps = h->steps;
  out->rebuilds = h->rebuilds;
  out->pair_touch = (int64_t)s[sqdk::kStatPairTouch];
  out->wall_touch = (int64_t)s[sqdk::kStatWallTouch];
  out->newton_last = (int64_t)s[sqdk::kStatNewton];
  out->newton = (int64_t)tot[sqdk::kStatNewton];
  out->checks = (int64_t)tot[sqdk::kStatChecked];
  out->fresh = (int64_t)tot[sqdk::kStatFresh];
  out->fresh_iters = (int64_t)tot[sqdk::kStatFreshIters];
  out->fallbacks = (int64_t)tot[sqdk::kStatFallback];
  out->run_ms = h->run_ms;
  return DEM_OK;
}

}  // extern "C"
