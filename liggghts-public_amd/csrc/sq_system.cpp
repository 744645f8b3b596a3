// Host side of the superquadric contact-point solve and its warm-start
// network (C ABI declared in include/liggghts_amd.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "../../include/liggghts_amd.h"
#include "sq_kernels.h"

struct dem_sq_handle {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  std::vector<void *> pair_allocs;  // per pair set
  int64_t n = 0;
  double *pose_a = nullptr, *pose_b = nullptr, *init = nullptr, *point = nullptr, *fi = nullptr,
         *fj = nullptr;
  int *iters = nullptr, *flags = nullptr;
  float *X = nullptr, *Y = nullptr, *T = nullptr;  // network inputs / outputs / targets of the pairs
  bool have_fresh = false;                         // point holds a fresh (plain) solve
  // contact detection buffers, allocated with the pair set on first use
  bool chk_ready = false;
  double *hist = nullptr, *vel = nullptr, *mass = nullptr, *en = nullptr, *deltan = nullptr, *cp_i = nullptr,
         *cp_j = nullptr;
  int *contact = nullptr;
  double solve_ms = 0.0, net_ms = 0.0;
  hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
  // network
  std::vector<void *> net_allocs;
  sqk::NetDev net{};
  float *grad = nullptr, *part = nullptr, *loss_part = nullptr, *loss_hist = nullptr;
  int64_t *idx = nullptr;
  int part_cap = 0;
  int64_t idx_cap = 0, loss_cap = 0;
  int64_t adam_t = 0;
  bool net_ready = false;
  // this device's copy of the allocation-free gradient's workspace
  float *scr_part = nullptr, *scr_loss = nullptr;
};

namespace {

std::string g_sq_create_error;

int sq_fail(dem_sq_handle *h, int code, const std::string &msg) {
  if (h)
    h->err = msg;
  else
    g_sq_create_error = msg;
  return code;
}

#define SQCHK(h, call)                                                              \
  do {                                                                              \
    hipError_t e_ = (call);                                                         \
    if (e_ != hipSuccess) return sq_fail(h, DEM_ERR_HIP, hipGetErrorString(e_)); \
  } while (0)

int alloc(dem_sq_handle *h, std::vector<void *> &list, void **p, size_t bytes) {
  *p = nullptr;
  if (bytes == 0) bytes = 16;
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) return sq_fail(h, DEM_ERR_HIP, std::string("hipMalloc: ") + hipGetErrorString(e));
  list.push_back(*p);
  return DEM_OK;
}

void free_list(std::vector<void *> &list) {
  for (void *p : list) hipFree(p);
  list.clear();
}

int ensure_net(dem_sq_handle *h) {
  if (!h->net_ready) return sq_fail(h, DEM_ERR_STATE, "network not initialised (dem_sq_net_init)");
  return DEM_OK;
}

// batch workspace: gradient partials, loss partials and row indices
int ensure_batch(dem_sq_handle *h, int64_t batch) {
  const int nb = sqk::grad_blocks(batch);
  int rc;
  if (nb > h->part_cap) {
    if ((rc = alloc(h, h->net_allocs, (void **)&h->part, sizeof(float) * (size_t)nb * sqk::kNumParams)))
      return rc;
    if ((rc = alloc(h, h->net_allocs, (void **)&h->loss_part, sizeof(float) * (size_t)nb))) return rc;
    h->part_cap = nb;
  }
  if (batch > h->idx_cap) {
    if ((rc = alloc(h, h->net_allocs, (void **)&h->idx, sizeof(int64_t) * (size_t)batch))) return rc;
    h->idx_cap = batch;
  }
  return DEM_OK;
}

}  // namespace

namespace sqk {
// the trained network of a solver handle, for the superquadric DEM step's
// warm start (null when the handle has no network)
const NetDev *handle_net(const dem_sq_handle *h, int *device) {
  if (!h || !h->net_ready) return nullptr;
  if (device) *device = h->device;
  return &h->net;
}
}  // namespace sqk

extern "C" {

const char *dem_sq_last_error(const dem_sq_handle *h) { return h ? h->err.c_str() : g_sq_create_error.c_str(); }

int dem_sq_create(int device, dem_sq_handle **out) {
  if (!out) return sq_fail(nullptr, DEM_ERR_ARG, "null argument");
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev)
    return sq_fail(nullptr, DEM_ERR_NODEV, "no HIP device available");
  if (hipSetDevice(device) != hipSuccess) return sq_fail(nullptr, DEM_ERR_NODEV, "hipSetDevice failed");
  dem_sq_handle *h = new dem_sq_handle();
  h->device = device;
  if (hipStreamCreate(&h->stream) != hipSuccess) {
    delete h;
    return sq_fail(nullptr, DEM_ERR_HIP, "hipStreamCreate failed");
  }
  for (auto &e : h->ev)
    if (hipEventCreate(&e) != hipSuccess) {
      delete h;
      return sq_fail(nullptr, DEM_ERR_HIP, "hipEventCreate failed");
    }
  const hipError_t se = sqk::grad_scratch(device, &h->scr_part, &h->scr_loss);
  if (se != hipSuccess) {
    dem_sq_destroy(h);
    return sq_fail(nullptr, DEM_ERR_HIP, std::string("hipGetSymbolAddress: ") + hipGetErrorString(se));
  }
  *out = h;
  return DEM_OK;
}

int dem_sq_destroy(dem_sq_handle *h) {
  if (!h) return DEM_OK;
  if (h->stream) hipStreamSynchronize(h->stream);
  free_list(h->pair_allocs);
  free_list(h->net_allocs);
  for (auto &e : h->ev)
    if (e) hipEventDestroy(e);
  if (h->stream) hipStreamDestroy(h->stream);
  delete h;
  return DEM_OK;
}

int dem_sq_set_pairs(dem_sq_handle *h, int64_t n, const double *pose_a, const double *pose_b) {
  if (!h) return DEM_ERR_ARG;
  if (n <= 0 || n > (int64_t)1 << 31 || !pose_a || !pose_b) return sq_fail(h, DEM_ERR_ARG, "bad pose arrays");
  for (int64_t k = 0; k < n; k++)
    for (const double *p : {pose_a + 12 * k, pose_b + 12 * k}) {
      for (int c = 0; c < 12; c++)
        if (!std::isfinite(p[c])) return sq_fail(h, DEM_ERR_ARG, "non-finite pose");
      if (!(p[7] > 0 && p[8] > 0 && p[9] > 0)) return sq_fail(h, DEM_ERR_ARG, "half-axes must be > 0");
      if (!(p[10] >= 2.0 && p[11] >= 2.0)) return sq_fail(h, DEM_ERR_ARG, "blockiness must be >= 2");
      const double qn = p[3] * p[3] + p[4] * p[4] + p[5] * p[5] + p[6] * p[6];
      if (std::fabs(qn - 1.0) > 1e-6) return sq_fail(h, DEM_ERR_ARG, "quaternions must be unit");
    }
  SQCHK(h, hipSetDevice(h->device));
  SQCHK(h, hipStreamSynchronize(h->stream));
  free_list(h->pair_allocs);
  h->n = n;
  h->have_fresh = false;
  h->chk_ready = false;
  const size_t N = (size_t)n;
  int rc;
  auto A = [&](void **p, size_t b) { return alloc(h, h->pair_allocs, p, b); };
  if ((rc = A((void **)&h->pose_a, sizeof(double) * 12 * N)) || (rc = A((void **)&h->pose_b, sizeof(double) * 12 * N)) ||
      (rc = A((void **)&h->init, sizeof(double) * 3 * N)) || (rc = A((void **)&h->point, sizeof(double) * 3 * N)) ||
      (rc = A((void **)&h->fi, sizeof(double) * N)) || (rc = A((void **)&h->fj, sizeof(double) * N)) ||
      (rc = A((void **)&h->iters, sizeof(int) * N)) || (rc = A((void **)&h->flags, sizeof(int) * N)) ||
      (rc = A((void **)&h->X, sizeof(float) * sqk::kIn * N)) || (rc = A((void **)&h->Y, sizeof(float) * sqk::kOut * N)) ||
      (rc = A((void **)&h->T, sizeof(float) * sqk::kOut * N)))
    return rc;
  SQCHK(h, hipMemcpy(h->pose_a, pose_a, sizeof(double) * 12 * N, hipMemcpyHostToDevice));
  SQCHK(h, hipMemcpy(h->pose_b, pose_b, sizeof(double) * 12 * N, hipMemcpyHostToDevice));
  SQCHK(h, sqk::launch_features(h->pose_a, h->pose_b, n, h->X, h->stream));
  SQCHK(h, hipStreamSynchronize(h->stream));
  return DEM_OK;
}

int dem_sq_solve(dem_sq_handle *h, int start, const double *init, int polish) {
  if (!h) return DEM_ERR_ARG;
  if (h->n <= 0) return sq_fail(h, DEM_ERR_STATE, "no pose pairs");
  if (start != DEM_SQ_FRESH && start != DEM_SQ_GIVEN && start != DEM_SQ_NETWORK)
    return sq_fail(h, DEM_ERR_ARG, "unknown start");
  if (start == DEM_SQ_GIVEN && !init) return sq_fail(h, DEM_ERR_ARG, "start points missing");
  if (polish < 0 || polish > 50) return sq_fail(h, DEM_ERR_ARG, "polish steps out of range");
  if (start == DEM_SQ_NETWORK) {
    int rc = ensure_net(h);
    if (rc) return rc;
  }
  SQCHK(h, hipSetDevice(h->device));
  if (start == DEM_SQ_GIVEN)
    SQCHK(h, hipMemcpy(h->init, init, sizeof(double) * 3 * (size_t)h->n, hipMemcpyHostToDevice));
  SQCHK(h, hipEventRecord(h->ev[0], h->stream));
  if (start == DEM_SQ_NETWORK) {
    SQCHK(h, sqk::launch_mlp_forward(h->net, h->X, h->n, h->Y, h->stream));
    SQCHK(h, sqk::launch_to_point(h->pose_a, h->pose_b, h->n, h->Y, h->init, h->stream));
  }
  SQCHK(h, hipEventRecord(h->ev[1], h->stream));
  sqk::SolveArgs a{};
  a.pose_a = h->pose_a;
  a.pose_b = h->pose_b;
  a.init = h->init;
  a.n = h->n;
  a.mode = start == DEM_SQ_FRESH ? 0 : 1;
  a.max_iter = 100000;  // Niter of calc_contact_point (math_extra_liggghts_superquadric.cpp:856)
  a.polish = polish;
  a.point = h->point;
  a.fi = h->fi;
  a.fj = h->fj;
  a.iters = h->iters;
  a.fail = h->flags;
  SQCHK(h, sqk::launch_solve(a, h->stream));
  SQCHK(h, hipEventRecord(h->ev[2], h->stream));
  SQCHK(h, hipEventSynchronize(h->ev[2]));
  float ms0 = 0.f, ms1 = 0.f;
  SQCHK(h, hipEventElapsedTime(&ms0, h->ev[0], h->ev[1]));
  SQCHK(h, hipEventElapsedTime(&ms1, h->ev[1], h->ev[2]));
  h->net_ms = ms0;
  h->solve_ms = ms1;
  h->have_fresh = start == DEM_SQ_FRESH && polish == 0;
  return DEM_OK;
}

int dem_sq_check(dem_sq_handle *h, int start, double dt, const double *vel, const double *mass, double *hist,
                 int *contact, double *en, double *deltan, double *cp_i, double *cp_j) {
  if (!h) return DEM_ERR_ARG;
  if (h->n <= 0) return sq_fail(h, DEM_ERR_STATE, "no pose pairs");
  if (start != DEM_SQ_FRESH && start != DEM_SQ_NETWORK)
    return sq_fail(h, DEM_ERR_ARG, "contact detection starts fresh contacts from the homotopy or the network");
  if (!hist) return sq_fail(h, DEM_ERR_ARG, "contact history missing");
  if (!vel != !mass) return sq_fail(h, DEM_ERR_ARG, "velocities and masses go together");
  if (!std::isfinite(dt) || dt < 0.0) return sq_fail(h, DEM_ERR_ARG, "bad time step");
  const size_t N = (size_t)h->n;
  for (size_t k = 0; k < N; k++) {
    const double *q = hist + sqk::kHist * k;
    for (int c = 0; c < sqk::kHist; c++)
      if (!std::isfinite(q[c])) return sq_fail(h, DEM_ERR_ARG, "non-finite contact history");
    if (!(q[0] == 0.0 || q[0] == 1.0 || q[0] == 2.0)) return sq_fail(h, DEM_ERR_ARG, "contact state must be 0, 1 or 2");
    if (!(q[1] >= 0.0 && q[1] < 15.0)) return sq_fail(h, DEM_ERR_ARG, "OBB inequality index must be in [0, 15)");
  }
  if (mass)
    for (size_t k = 0; k < 2 * N; k++)
      if (!(mass[k] > 0.0) || !std::isfinite(mass[k])) return sq_fail(h, DEM_ERR_ARG, "masses must be > 0");
  if (start == DEM_SQ_NETWORK) {
    int rc = ensure_net(h);
    if (rc) return rc;
  }
  SQCHK(h, hipSetDevice(h->device));
  if (!h->chk_ready) {
    int rc;
    auto A = [&](void **p, size_t b) { return alloc(h, h->pair_allocs, p, b); };
    if ((rc = A((void **)&h->hist, sizeof(double) * sqk::kHist * N)) || (rc = A((void **)&h->vel, sizeof(double) * 12 * N)) ||
        (rc = A((void **)&h->mass, sizeof(double) * 2 * N)) || (rc = A((void **)&h->en, sizeof(double) * 3 * N)) ||
        (rc = A((void **)&h->deltan, sizeof(double) * N)) || (rc = A((void **)&h->cp_i, sizeof(double) * 3 * N)) ||
        (rc = A((void **)&h->cp_j, sizeof(double) * 3 * N)) || (rc = A((void **)&h->contact, sizeof(int) * N)))
      return rc;
    h->chk_ready = true;
  }
  SQCHK(h, hipMemcpy(h->hist, hist, sizeof(double) * sqk::kHist * N, hipMemcpyHostToDevice));
  if (vel) {
    SQCHK(h, hipMemcpy(h->vel, vel, sizeof(double) * 12 * N, hipMemcpyHostToDevice));
    SQCHK(h, hipMemcpy(h->mass, mass, sizeof(double) * 2 * N, hipMemcpyHostToDevice));
  }
  SQCHK(h, hipEventRecord(h->ev[0], h->stream));
  if (start == DEM_SQ_NETWORK) {
    SQCHK(h, sqk::launch_mlp_forward(h->net, h->X, h->n, h->Y, h->stream));
    SQCHK(h, sqk::launch_to_point(h->pose_a, h->pose_b, h->n, h->Y, h->init, h->stream));
  }
  SQCHK(h, hipEventRecord(h->ev[1], h->stream));
  sqk::CheckArgs a{};
  a.pose_a = h->pose_a;
  a.pose_b = h->pose_b;
  a.init = start == DEM_SQ_NETWORK ? h->init : nullptr;
  a.vel = vel ? h->vel : nullptr;
  a.mass = vel ? h->mass : nullptr;
  a.dt = dt;
  a.n = h->n;
  a.max_iter = 100000;  // Niter of calc_contact_point (math_extra_liggghts_superquadric.cpp:856)
  a.hist = h->hist;
  a.contact = h->contact;
  a.en = h->en;
  a.deltan = h->deltan;
  a.cp_i = h->cp_i;
  a.cp_j = h->cp_j;
  a.point = h->point;
  a.fi = h->fi;
  a.fj = h->fj;
  a.iters = h->iters;
  a.flags = h->flags;
  SQCHK(h, sqk::launch_check(a, h->stream));
  SQCHK(h, hipEventRecord(h->ev[2], h->stream));
  SQCHK(h, hipEventSynchronize(h->ev[2]));
  float ms0 = 0.f, ms1 = 0.f;
  SQCHK(h, hipEventElapsedTime(&ms0, h->ev[0], h->ev[1]));
  SQCHK(h, hipEventElapsedTime(&ms1, h->ev[1], h->ev[2]));
  h->net_ms = ms0;
  h->solve_ms = ms1;
  h->have_fresh = false;
  SQCHK(h, hipMemcpy(hist, h->hist, sizeof(double) * sqk::kHist * N, hipMemcpyDeviceToHost));
  if (contact) SQCHK(h, hipMemcpy(contact, h->contact, sizeof(int) * N, hipMemcpyDeviceToHost));
  if (en) SQCHK(h, hipMemcpy(en, h->en, sizeof(double) * 3 * N, hipMemcpyDeviceToHost));
  if (deltan) SQCHK(h, hipMemcpy(deltan, h->deltan, sizeof(double) * N, hipMemcpyDeviceToHost));
  if (cp_i) SQCHK(h, hipMemcpy(cp_i, h->cp_i, sizeof(double) * 3 * N, hipMemcpyDeviceToHost));
  if (cp_j) SQCHK(h, hipMemcpy(cp_j, h->cp_j, sizeof(double) * 3 * N, hipMemcpyDeviceToHost));
  return DEM_OK;
}

int dem_sq_get_results(dem_sq_handle *h, double *point, double *fi, double *fj, int *iters, int *flags,
                       double *solve_ms, double *net_ms) {
  if (!h) return DEM_ERR_ARG;
  if (h->n <= 0) return sq_fail(h, DEM_ERR_STATE, "no pose pairs");
  SQCHK(h, hipStreamSynchronize(h->stream));
  const size_t N = (size_t)h->n;
  if (point) SQCHK(h, hipMemcpy(point, h->point, sizeof(double) * 3 * N, hipMemcpyDeviceToHost));
  if (fi) SQCHK(h, hipMemcpy(fi, h->fi, sizeof(double) * N, hipMemcpyDeviceToHost));
  if (fj) SQCHK(h, hipMemcpy(fj, h->fj, sizeof(double) * N, hipMemcpyDeviceToHost));
  if (iters) SQCHK(h, hipMemcpy(iters, h->iters, sizeof(int) * N, hipMemcpyDeviceToHost));
  if (flags) SQCHK(h, hipMemcpy(flags, h->flags, sizeof(int) * N, hipMemcpyDeviceToHost));
  if (solve_ms) *solve_ms = h->solve_ms;
  if (net_ms) *net_ms = h->net_ms;
  return DEM_OK;
}

int dem_sq_get_features(dem_sq_handle *h, float *X) {
  if (!h || !X) return DEM_ERR_ARG;
  if (h->n <= 0) return sq_fail(h, DEM_ERR_STATE, "no pose pairs");
  SQCHK(h, hipMemcpy(X, h->X, sizeof(float) * sqk::kIn * (size_t)h->n, hipMemcpyDeviceToHost));
  return DEM_OK;
}

int dem_sq_net_init(dem_sq_handle *h, uint64_t seed) {
  if (!h) return DEM_ERR_ARG;
  SQCHK(h, hipSetDevice(h->device));
  if (!h->net_ready) {
    int rc;
    auto A = [&](void **p, size_t b) { return alloc(h, h->net_allocs, p, b); };
    sqk::NetDev &n = h->net;
    if ((rc = A((void **)&n.params, sizeof(float) * sqk::kNumParams)) || (rc = A((void **)&n.m, sizeof(float) * sqk::kNumParams)) ||
        (rc = A((void **)&n.v, sizeof(float) * sqk::kNumParams)) ||
        (rc = A((void **)&n.w1b, 2 * sqk::kInPad * sqk::kHid)) || (rc = A((void **)&n.w2b, 2 * sqk::kHid * sqk::kHid)) ||
        (rc = A((void **)&n.w3b, 2 * sqk::kHid * sqk::kOutPad)) || (rc = A((void **)&n.b1, sizeof(float) * sqk::kHid)) ||
        (rc = A((void **)&n.b2, sizeof(float) * sqk::kHid)) || (rc = A((void **)&n.b3, sizeof(float) * sqk::kOutPad)) ||
        (rc = A((void **)&h->grad, sizeof(float) * sqk::kNumParams)))
      return rc;
  }
  // Glorot-uniform weights, zero biases
  std::mt19937_64 rng(seed);
  std::vector<float> p(sqk::kNumParams, 0.f);
  auto fill = [&](int off, int fan_in, int fan_out) {
    const double lim = std::sqrt(6.0 / (fan_in + fan_out));
    std::uniform_real_distribution<double> U(-lim, lim);
    for (int k = 0; k < fan_in * fan_out; k++) p[off + k] = (float)U(rng);
  };
  fill(sqk::kOffW1, sqk::kIn, sqk::kHid);
  fill(sqk::kOffW2, sqk::kHid, sqk::kHid);
  fill(sqk::kOffW3, sqk::kHid, sqk::kOut);
  h->net_ready = true;
  return dem_sq_net_set_params(h, p.data());
}

int dem_sq_net_set_params(dem_sq_handle *h, const float *params) {
  if (!h || !params) return DEM_ERR_ARG;
  if (!h->net_ready) return sq_fail(h, DEM_ERR_STATE, "network not initialised (dem_sq_net_init)");
  const sqk::NetDev &n = h->net;
  SQCHK(h, hipStreamSynchronize(h->stream));
  SQCHK(h, hipMemcpy(n.params, params, sizeof(float) * sqk::kNumParams, hipMemcpyHostToDevice));
  SQCHK(h, hipMemset(n.m, 0, sizeof(float) * sqk::kNumParams));
  SQCHK(h, hipMemset(n.v, 0, sizeof(float) * sqk::kNumParams));
  SQCHK(h, hipMemset(n.w1b, 0, 2 * sqk::kInPad * sqk::kHid));
  SQCHK(h, hipMemset(n.w3b, 0, 2 * sqk::kHid * sqk::kOutPad));
  SQCHK(h, hipMemset(n.b3, 0, sizeof(float) * sqk::kOutPad));
  SQCHK(h, sqk::launch_mlp_pack(n, h->stream));
  SQCHK(h, hipStreamSynchronize(h->stream));
  h->adam_t = 0;
  return DEM_OK;
}

int dem_sq_net_get_params(dem_sq_handle *h, float *params) {
  if (!h || !params) return DEM_ERR_ARG;
  int rc = ensure_net(h);
  if (rc) return rc;
  SQCHK(h, hipStreamSynchronize(h->stream));
  SQCHK(h, hipMemcpy(params, h->net.params, sizeof(float) * sqk::kNumParams, hipMemcpyDeviceToHost));
  return DEM_OK;
}

int dem_sq_net_forward(dem_sq_handle *h, int64_t n, const float *X, float *Y) {
  if (!h || n <= 0 || !X || !Y) return DEM_ERR_ARG;
  int rc = ensure_net(h);
  if (rc) return rc;
  std::vector<void *> tmp;
  float *dX, *dY;
  if ((rc = alloc(h, tmp, (void **)&dX, sizeof(float) * sqk::kIn * n)) ||
      (rc = alloc(h, tmp, (void **)&dY, sizeof(float) * sqk::kOut * n))) {
    free_list(tmp);
    return rc;
  }
  hipError_t e = hipMemcpy(dX, X, sizeof(float) * sqk::kIn * n, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = sqk::launch_mlp_forward(h->net, dX, n, dY, h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  if (e == hipSuccess) e = hipMemcpy(Y, dY, sizeof(float) * sqk::kOut * n, hipMemcpyDeviceToHost);
  free_list(tmp);
  if (e != hipSuccess) return sq_fail(h, DEM_ERR_HIP, hipGetErrorString(e));
  return DEM_OK;
}

int dem_sq_net_gradient(dem_sq_handle *h, int64_t n, const float *X, const float *T, float *grad,
                        float *loss) {
  if (!h || n <= 0 || !X || !T || !grad) return DEM_ERR_ARG;
  int rc = ensure_net(h);
  if (rc) return rc;
  if ((rc = ensure_batch(h, n))) return rc;
  std::vector<void *> tmp;
  float *dX, *dT, *dL;
  if ((rc = alloc(h, tmp, (void **)&dX, sizeof(float) * sqk::kIn * n)) ||
      (rc = alloc(h, tmp, (void **)&dT, sizeof(float) * sqk::kOut * n)) ||
      (rc = alloc(h, tmp, (void **)&dL, sizeof(float)))) {
    free_list(tmp);
    return rc;
  }
  std::vector<int64_t> rows((size_t)n);
  for (int64_t k = 0; k < n; k++) rows[(size_t)k] = k;
  sqk::AdamArgs ad{};
  ad.apply = 0;
  ad.loss_part = h->loss_part;
  ad.loss_out = dL;
  ad.batch = n;
  hipError_t e = hipMemcpy(dX, X, sizeof(float) * sqk::kIn * n, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(dT, T, sizeof(float) * sqk::kOut * n, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(h->idx, rows.data(), sizeof(int64_t) * n, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = sqk::launch_mlp_grad(h->net, dX, dT, h->idx, n, h->part, h->loss_part, h->stream);
  if (e == hipSuccess) e = sqk::launch_mlp_adam(h->net, h->part, sqk::grad_blocks(n), h->grad, ad, h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  if (e == hipSuccess) e = hipMemcpy(grad, h->grad, sizeof(float) * sqk::kNumParams, hipMemcpyDeviceToHost);
  float l = 0.f;
  if (e == hipSuccess) e = hipMemcpy(&l, dL, sizeof(float), hipMemcpyDeviceToHost);
  free_list(tmp);
  if (e != hipSuccess) return sq_fail(h, DEM_ERR_HIP, hipGetErrorString(e));
  if (loss) *loss = l;
  return DEM_OK;
}

static_assert(sqk::kGradScratchMaxBatch == DEM_SQ_GRAD_NOALLOC_MAX, "workspace limit and the header disagree");

int dem_sq_net_gradient_noalloc(dem_sq_handle *h, int64_t n, const float *X, const float *T, float *grad,
                                float *loss) {
  if (!h) return DEM_ERR_ARG;
  if (!X || !T || !grad) return sq_fail(h, DEM_ERR_ARG, "null device pointer");
  if (n <= 0 || n > sqk::kGradScratchMaxBatch)
    return sq_fail(h, DEM_ERR_ARG,
                   "batch of " + std::to_string(n) + " samples outside 1.." +
                       std::to_string(sqk::kGradScratchMaxBatch) + " (the fixed workspace's limit)");
  int rc = ensure_net(h);
  if (rc) return rc;
  if (!h->scr_part) return sq_fail(h, DEM_ERR_STATE, "workspace address not resolved");
  SQCHK(h, hipSetDevice(h->device));
  sqk::AdamArgs ad{};
  ad.apply = 0;
  ad.loss_part = h->scr_loss;
  ad.loss_out = loss;
  ad.batch = n;
  SQCHK(h, sqk::launch_mlp_grad(h->net, X, T, nullptr, n, h->scr_part, h->scr_loss, h->stream));
  SQCHK(h, sqk::launch_mlp_adam(h->net, h->scr_part, sqk::grad_blocks(n), grad, ad, h->stream));
  SQCHK(h, hipStreamSynchronize(h->stream));
  return DEM_OK;
}

int dem_sq_net_train(dem_sq_handle *h, int64_t steps, int64_t batch, float lr, uint64_t seed,
                     float *loss_curve, double *train_ms) {
  if (!h || steps <= 0 || batch <= 0 || !(lr > 0.f)) return DEM_ERR_ARG;
  int rc = ensure_net(h);
  if (rc) return rc;
  if (!h->have_fresh) return sq_fail(h, DEM_ERR_STATE, "training needs a fresh (plain) solve of the pairs");
  if ((rc = ensure_batch(h, batch))) return rc;
  if (steps > h->loss_cap) {
    if ((rc = alloc(h, h->net_allocs, (void **)&h->loss_hist, sizeof(float) * (size_t)steps))) return rc;
    h->loss_cap = steps;
  }
  SQCHK(h, hipSetDevice(h->device));
  // targets: the plain solve's contact points in A's body frame
  SQCHK(h, sqk::launch_to_target(h->pose_a, h->pose_b, h->n, h->point, h->T, h->stream));
  const int nb = sqk::grad_blocks(batch);
  sqk::AdamArgs ad{};
  ad.apply = 1;
  ad.lr = lr;
  ad.beta1 = 0.9f;
  ad.beta2 = 0.999f;
  ad.eps = 1e-8f;
  ad.loss_part = h->loss_part;
  ad.batch = batch;
  SQCHK(h, hipEventRecord(h->ev[0], h->stream));
  for (int64_t s = 0; s < steps; s++) {
    h->adam_t++;
    ad.bc1 = (float)(1.0 - std::pow(0.9, (double)h->adam_t));
    ad.bc2 = (float)(1.0 - std::pow(0.999, (double)h->adam_t));
    ad.loss_out = h->loss_hist + s;
    SQCHK(h, sqk::launch_sample(h->idx, batch, h->n, seed + 0x632be59bd9b4e019ull * (uint64_t)s, h->stream));
    SQCHK(h, sqk::launch_mlp_grad(h->net, h->X, h->T, h->idx, batch, h->part, h->loss_part, h->stream));
    SQCHK(h, sqk::launch_mlp_adam(h->net, h->part, nb, h->grad, ad, h->stream));
  }
  SQCHK(h, hipEventRecord(h->ev[1], h->stream));
  SQCHK(h, hipEventSynchronize(h->ev[1]));
  float ms = 0.f;
  SQCHK(h, hipEventElapsedTime(&ms, h->ev[0], h->ev[1]));
  if (train_ms) *train_ms = ms;
  if (loss_curve) SQCHK(h, hipMemcpy(loss_curve, h->loss_hist, sizeof(float) * (size_t)steps, hipMemcpyDeviceToHost));
  return DEM_OK;
}

}  // extern "C"
