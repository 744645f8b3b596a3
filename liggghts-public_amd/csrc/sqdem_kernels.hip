// Superquadric DEM step on the GPU, FP64: the per-timestep path of
// atom_style superquadric with pair_style gran model hertz tangential history
// surface superquadric, fix wall/gran ... primitive, fix gravity and
// fix nve/superquadric (integration_scheme 1).
//
// Data: per-particle arrays in the caller's order (the expensive part is the
// per-pair Newton solve, not the particle streams, so no reordering); a half
// pair list (i < j) in CSR form by i with j ascending, and the same pairs
// listed by j. The pair kernel runs one thread per pair and writes a force
// record (force on i, torque on i, torque on j); the gather kernel sums each
// particle's records in a fixed order (pairs where it is j, then pairs where
// it is i, each by pair index), so results do not depend on thread timing.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "dem_device.h"
#include "sq_device.h"
#include "sqdem_kernels.h"

namespace sqdk {

constexpr int kBlk = 256;

static inline int nblk(int64_t n, int b) { return (int)((n + b - 1) / b); }

// MathExtraLiggghtsNonspherical::quat_to_mat (math_extra_liggghts_nonspherical.h:511-534)
__device__ __forceinline__ void rot_of(const double *q, double *m) {
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

__device__ __forceinline__ void mv(const double *m, const double *v, double *r) {
  r[0] = m[0] * v[0] + m[1] * v[1] + m[2] * v[2];
  r[1] = m[3] * v[0] + m[4] * v[1] + m[5] * v[2];
  r[2] = m[6] * v[0] + m[7] * v[1] + m[8] * v[2];
}

__device__ __forceinline__ void tmv(const double *m, const double *v, double *r) {
  r[0] = m[0] * v[0] + m[3] * v[1] + m[6] * v[2];
  r[1] = m[1] * v[0] + m[4] * v[1] + m[7] * v[2];
  r[2] = m[2] * v[0] + m[5] * v[1] + m[8] * v[2];
}

__device__ __forceinline__ void cross(const double *a, const double *b, double *r) {
  r[0] = a[1] * b[2] - a[2] * b[1];
  r[1] = a[2] * b[0] - a[0] * b[2];
  r[2] = a[0] * b[1] - a[1] * b[0];
}

// MathExtra::quatvec / invquatvec (math_extra.h:584-616)
__device__ __forceinline__ void quatvec(const double *a, const double *b, double *c) {
  c[0] = -a[1] * b[0] - a[2] * b[1] - a[3] * b[2];
  c[1] = a[0] * b[0] + a[2] * b[2] - a[3] * b[1];
  c[2] = a[0] * b[1] + a[3] * b[0] - a[1] * b[2];
  c[3] = a[0] * b[2] + a[1] * b[1] - a[2] * b[0];
}

__device__ __forceinline__ void invquatvec(const double *a, const double *b, double *c) {
  c[0] = -a[1] * b[0] + a[0] * b[1] + a[3] * b[2] - a[2] * b[3];
  c[1] = -a[2] * b[0] - a[3] * b[1] + a[0] * b[2] + a[1] * b[3];
  c[2] = -a[3] * b[0] + a[2] * b[1] - a[1] * b[2] + a[0] * b[3];
}

// no_squish_rotate (math_extra_liggghts_nonspherical.cpp:175-216)
__device__ __forceinline__ void no_squish_rotate(int k, double *p, double *q, const double *inertia, double dt) {
  double kp[4], kq[4];
  if (k == 1) {
    kq[0] = -q[1]; kp[0] = -p[1];
    kq[1] = q[0];  kp[1] = p[0];
    kq[2] = q[3];  kp[2] = p[3];
    kq[3] = -q[2]; kp[3] = -p[2];
  } else if (k == 2) {
    kq[0] = -q[2]; kp[0] = -p[2];
    kq[1] = -q[3]; kp[1] = -p[3];
    kq[2] = q[0];  kp[2] = p[0];
    kq[3] = q[1];  kp[3] = p[1];
  } else {
    kq[0] = -q[3]; kp[0] = -p[3];
    kq[1] = q[2];  kp[1] = p[2];
    kq[2] = -q[1]; kp[2] = -p[1];
    kq[3] = q[0];  kp[3] = p[0];
  }
  double phi = p[0] * kq[0] + p[1] * kq[1] + p[2] * kq[2] + p[3] * kq[3];
  phi /= 4.0 * inertia[k - 1];
  const double c_phi = cos(dt * phi), s_phi = sin(dt * phi);
  for (int m = 0; m < 4; m++) p[m] = c_phi * p[m] + s_phi * kp[m];
  for (int m = 0; m < 4; m++) q[m] = c_phi * q[m] + s_phi * kq[m];
}

// FixNVEAsphereBase::initial_integrate, integration_scheme 1
// (fix_nve_asphere_base.cpp:355-460), then the displacement test of
// Neighbor::check_distance (neighbor.cpp:1427-1461)
__global__ __launch_bounds__(kBlk) void k_initial(Dev d, int *flag) {
  const int i = blockIdx.x * kBlk + threadIdx.x;
  if (i >= d.n) return;
  const double dtv = d.dt, dtf = 0.5 * d.dt, dtf2 = 2.0 * dtf, dtq = 0.5 * dtv;
  double x[3], v[3], L[3], q[4], f[3], t[3], I[3];
  for (int k = 0; k < 3; k++) {
    x[k] = d.x[3 * i + k];
    v[k] = d.v[3 * i + k];
    L[k] = d.L[3 * i + k];
    f[k] = d.f[3 * i + k];
    t[k] = d.t[3 * i + k];
    I[k] = d.I[3 * i + k];
  }
  for (int k = 0; k < 4; k++) q[k] = d.q[4 * i + k];
  double R[9], mbody[3], conjqm[4], tbody[3], fquat[4], wbody[3], w[3];
  rot_of(q, R);
  const double dtfm = dtf / d.mass[i];
  for (int k = 0; k < 3; k++) v[k] += dtfm * f[k];
  for (int k = 0; k < 3; k++) x[k] += dtv * v[k];
  tmv(R, L, mbody);
  quatvec(q, mbody, conjqm);
  for (int k = 0; k < 4; k++) conjqm[k] *= 2.0;
  tmv(R, t, tbody);
  quatvec(q, tbody, fquat);
  for (int k = 0; k < 4; k++) conjqm[k] += dtf2 * fquat[k];
  no_squish_rotate(3, conjqm, q, I, dtq);
  no_squish_rotate(2, conjqm, q, I, dtq);
  no_squish_rotate(1, conjqm, q, I, dtv);
  no_squish_rotate(2, conjqm, q, I, dtq);
  no_squish_rotate(3, conjqm, q, I, dtq);
  const double norm = 1.0 / sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
  for (int k = 0; k < 4; k++) q[k] *= norm;
  rot_of(q, R);
  invquatvec(q, conjqm, mbody);
  for (int k = 0; k < 3; k++) mbody[k] *= 0.5;
  for (int k = 0; k < 3; k++) wbody[k] = mbody[k] / I[k];
  mv(R, mbody, L);
  mv(R, wbody, w);
  for (int k = 0; k < 3; k++) {
    d.x[3 * i + k] = x[k];
    d.v[3 * i + k] = v[k];
    d.L[3 * i + k] = L[k];
    d.w[3 * i + k] = w[k];
  }
  for (int k = 0; k < 4; k++) d.q[4 * i + k] = q[k];
  const double trigger = 0.5 * d.skin;
  const double dx = x[0] - d.xhold[3 * i], dy = x[1] - d.xhold[3 * i + 1], dz = x[2] - d.xhold[3 * i + 2];
  if (dx * dx + dy * dy + dz * dz > trigger * trigger) *flag = 1;
}

// Domain::pbc (domain.cpp:530-620) on a rebuild step, keeping the shift
__global__ __launch_bounds__(kBlk) void k_wrap(Dev d) {
  const int i = blockIdx.x * kBlk + threadIdx.x;
  if (i >= d.n) return;
  for (int k = 0; k < 3; k++) {
    double xi = d.x[3 * i + k];
    const double x0 = xi;
    if (d.per[k]) {
      const double lo = d.lo[k], hi = d.hi[k], L = hi - lo;
      if (xi < lo) xi += L;
      if (xi >= hi) {
        xi -= L;
        if (xi < lo) xi = lo;
      }
    }
    d.x[3 * i + k] = xi;
    d.wrap[3 * i + k] = xi - x0;
    d.xhold[3 * i + k] = xi;
  }
}

__device__ __forceinline__ int cell_coord(const Dev &d, const Grid &g, double x, int k) {
  int c = (int)floor((x - d.lo[k]) / g.cs[k]);
  return c < 0 ? 0 : (c >= g.nc[k] ? g.nc[k] - 1 : c);
}

__global__ __launch_bounds__(kBlk) void k_zero_int(int *p, int n) {
  const int i = blockIdx.x * kBlk + threadIdx.x;
  if (i < n) p[i] = 0;
}

__global__ __launch_bounds__(kBlk) void k_bin_count(Dev d, Grid g) {
  const int i = blockIdx.x * kBlk + threadIdx.x;
  if (i >= d.n) return;
  const int cx = cell_coord(d, g, d.x[3 * i], 0), cy = cell_coord(d, g, d.x[3 * i + 1], 1),
            cz = cell_coord(d, g, d.x[3 * i + 2], 2);
  const int c = (cz * g.nc[1] + cy) * g.nc[0] + cx;
  g.cell_of[i] = c;
  atomicAdd(&g.count[c], 1);
}

__global__ __launch_bounds__(kBlk) void k_bin_fill(Dev d, Grid g) {
  const int i = blockIdx.x * kBlk + threadIdx.x;
  if (i >= d.n) return;
  const int c = g.cell_of[i];
  g.list[g.start[c] + atomicAdd(&g.fill[c], 1)] = i;
}

// insertion sort of a short segment of ints, ascending
__device__ __forceinline__ void sort_segment(int *a, int n) {
  for (int k = 1; k < n; k++) {
    const int v = a[k];
    int m = k - 1;
    while (m >= 0 && a[m] > v) {
      a[m + 1] = a[m];
      m--;
    }
    a[m + 1] = v;
  }
}

__global__ __launch_bounds__(kBlk) void k_bin_sort(Grid g) {
  const int c = blockIdx.x * kBlk + threadIdx.x;
  if (c >= g.ncell) return;
  sort_segment(g.list + g.start[c], g.start[c + 1] - g.start[c]);
}

// the periodic image of j nearest to i: x_j + k L, k in {-1, 0, 1}
__device__ __forceinline__ void image_of(const Dev &d, int i, int j, double *xj, double *del) {
  for (int k = 0; k < 3; k++) {
    double xjk = d.x[3 * j + k];
    const double xik = d.x[3 * i + k];
    if (d.per[k]) {
      const double dd = xik - xjk;
      if (dd > 0.5 * d.len[k])
        xjk = xjk + d.len[k];
      else if (dd < -0.5 * d.len[k])
        xjk = xjk - d.len[k];
    }
    xj[k] = xjk;
    del[k] = xik - xjk;
  }
}

// visits every j > i within the cutoff (rb_i + rb_j + skin), cells of the
// 27-cell stencil in a fixed order
template <typename F>
__device__ __forceinline__ void for_partners(const Dev &d, const Grid &g, int i, F &&fn) {
  const int c = g.cell_of[i];
  const int cx = c % g.nc[0], cy = (c / g.nc[0]) % g.nc[1], cz = c / (g.nc[0] * g.nc[1]);
  const double rbi = d.rb[i];
  for (int dz = -1; dz <= 1; dz++) {
    int z = cz + dz;
    if (d.per[2]) z = (z + g.nc[2]) % g.nc[2];
    else if (z < 0 || z >= g.nc[2]) continue;
    for (int dy = -1; dy <= 1; dy++) {
      int y = cy + dy;
      if (d.per[1]) y = (y + g.nc[1]) % g.nc[1];
      else if (y < 0 || y >= g.nc[1]) continue;
      for (int dx = -1; dx <= 1; dx++) {
        int x = cx + dx;
        if (d.per[0]) x = (x + g.nc[0]) % g.nc[0];
        else if (x < 0 || x >= g.nc[0]) continue;
        const int cc = (z * g.nc[1] + y) * g.nc[0] + x;
        for (int s = g.start[cc]; s < g.start[cc + 1]; s++) {
          const int j = g.list[s];
          if (j <= i) continue;
          double xj[3], del[3];
          image_of(d, i, j, xj, del);
          const double rsq = del[0] * del[0] + del[1] * del[1] + del[2] * del[2];
          const double cut = rbi + d.rb[j] + d.skin;
          if (rsq > cut * cut) continue;
          fn(j, rsq);
        }
      }
    }
  }
}

__global__ __launch_bounds__(kBlk) void k_pair_count(Dev d, Grid g, Build b) {
  const int i = blockIdx.x * kBlk + threadIdx.x;
  if (i >= d.n) return;
  int cnt = 0;
  for_partners(d, g, i, [&](int, double) { cnt++; });
  b.count[i] = cnt;
}

// fills row i, sorts it by j, and carries the history of pairs flagged as
// touching whose bounding spheres overlap (the contact point moved with i's
// periodic wrap)
__global__ __launch_bounds__(kBlk) void k_pair_fill(Dev d, Grid g, Build b) {
  const int i = blockIdx.x * kBlk + threadIdx.x;
  if (i >= d.n) return;
  const int f0 = b.first[i], f1 = b.first[i + 1];
  if (f1 > b.cap) {
    *b.overflow = 1;
    return;
  }
  int k = f0;
  for_partners(d, g, i, [&](int j, double) {
    if (k < f1) b.pj[k++] = j;
  });
  sort_segment(b.pj + f0, f1 - f0);
  const int o0 = b.ofirst ? b.ofirst[i] : 0, o1 = b.ofirst ? b.ofirst[i + 1] : 0;
  for (int e = f0; e < f1; e++) {
    const int j = b.pj[e];
    b.pi[e] = i;
    double *h = b.hist + (size_t)kHist * e;
    for (int m = 0; m < kHist; m++) h[m] = 0.0;
    int fl = 0;
    double xj[3], del[3];
    image_of(d, i, j, xj, del);
    const double rsq = del[0] * del[0] + del[1] * del[1] + del[2] * del[2];
    const double radsum = d.rb[i] + d.rb[j];
    if (rsq < radsum * radsum && o1 > o0) {
      int lo = o0, hi = o1 - 1;
      while (lo <= hi) {
        const int mid = (lo + hi) >> 1;
        const int jm = b.opj[mid];
        if (jm == j) {
          if (b.oflag[mid]) {
            const double *oh = b.ohist + (size_t)kHist * mid;
            for (int m = 0; m < kHist; m++) h[m] = oh[m];
            for (int m = 0; m < 3; m++) h[2 + m] += d.wrap[3 * i + m];
            fl = 1;
          }
          break;
        }
        if (jm < j)
          lo = mid + 1;
        else
          hi = mid - 1;
      }
    }
    b.flag[e] = fl;
  }
}

__global__ __launch_bounds__(kBlk) void k_rev_count(Build b, int np) {
  const int p = blockIdx.x * kBlk + threadIdx.x;
  if (p >= np) return;
  atomicAdd(&b.rcount[b.pj[p]], 1);
}

__global__ __launch_bounds__(kBlk) void k_rev_fill(Build b, int np) {
  const int p = blockIdx.x * kBlk + threadIdx.x;
  if (p >= np) return;
  const int j = b.pj[p];
  b.rlist[b.rfirst[j] + atomicAdd(&b.rfill[j], 1)] = p;
}

__global__ __launch_bounds__(kBlk) void k_rev_sort(Build b, int n) {
  const int j = blockIdx.x * kBlk + threadIdx.x;
  if (j >= n) return;
  sort_segment(b.rlist + b.rfirst[j], b.rfirst[j + 1] - b.rfirst[j]);
}

__global__ void k_zero_stats(unsigned long long *stats) {
  if (threadIdx.x < kNumStats) stats[threadIdx.x] = 0ull;
}

__global__ void k_accumulate(const unsigned long long *stats, unsigned long long *totals) {
  if (threadIdx.x < kNumStats) totals[threadIdx.x] += stats[threadIdx.x];
}

__device__ __forceinline__ void load_pose(const Dev &d, int i, const double *center, double *pose) {
  for (int k = 0; k < 3; k++) pose[k] = center[k];
  for (int k = 0; k < 4; k++) pose[3 + k] = d.q[4 * i + k];
  for (int k = 0; k < 3; k++) pose[7 + k] = d.shape[3 * i + k];
  pose[10] = d.blk[2 * i];
  pose[11] = d.blk[2 * i + 1];
}

__global__ __launch_bounds__(kBlk) void k_fresh_scan(Dev d, int *fresh_idx, int *fresh_count, double *pa,
                                                     double *pb) {
  const int64_t p = (int64_t)blockIdx.x * kBlk + threadIdx.x;
  if (p >= d.np) return;
  fresh_idx[p] = -1;
  const double *h = d.hist + (size_t)kHist * p;
  if ((int)(h[0] + 0.5) != 0) return;
  const int i = d.pi[p], j = d.pj[p];
  double xj[3], del[3];
  image_of(d, i, j, xj, del);
  const double rsq = del[0] * del[0] + del[1] * del[1] + del[2] * del[2];
  const double radsum = d.rb[i] + d.rb[j];
  if (!(rsq < radsum * radsum)) return;
  double poseA[12], poseB[12];
  load_pose(d, i, d.x + 3 * i, poseA);
  load_pose(d, j, xj, poseB);
  sqd::Sq A, B;
  sqd::init(A, poseA);
  sqd::init(B, poseB);
  int start = (int)(h[1] + 0.5);
  if (!sqd::obb_intersect(A, B, start)) return;
  const int slot = atomicAdd(fresh_count, 1);
  fresh_idx[p] = slot;
  for (int k = 0; k < 12; k++) {
    pa[12 * (size_t)slot + k] = poseA[k];
    pb[12 * (size_t)slot + k] = poseB[k];
  }
}

// Hertz normal force and Coulomb-capped tangential spring with contact-point
// lever arms: NormalModel<HERTZ> (normal_model_hertz.h:202-405) and
// TangentialModel<TANGENTIAL_HISTORY> (tangential_model_history.h:130-352)
// after surfacesIntersectNonSpherical (math_extra_liggghts_nonspherical.h:536-577)
__device__ __forceinline__ void contact_force(const Dev &d, int itype, int jtype, int is_wall, double reff,
                                              double meff, double deltan, const double *en, const double *vi,
                                              const double *vj, const double *wi, const double *wj,
                                              const double *xci, const double *xcj, int shearupdate,
                                              double *shear, double *dFi, double *dTi, double *dTj) {
  double vri[3], vrr[3], vrel[3];
  cross(wi, xci, vri);
  if (is_wall) {
    for (int k = 0; k < 3; k++) vrr[k] = vri[k];
  } else {
    double vrj[3];
    cross(wj, xcj, vrj);
    for (int k = 0; k < 3; k++) vrr[k] = vri[k] - vrj[k];
  }
  for (int k = 0; k < 3; k++) vrel[k] = vi[k] - vj[k] + vrr[k];
  const double vn = vrel[0] * en[0] + vrel[1] * en[1] + vrel[2] * en[2];
  const double vt[3] = {vrel[0] - vn * en[0], vrel[1] - vn * en[1], vrel[2] - vn * en[2]};
  const int tt = itype * d.t1 + jtype;
  const double Y = d.Y[tt], G = d.G[tt], Bt = d.B[tt];
  const double sqrtval = sqrt(reff * deltan);
  const double Sn = 2. * Y * sqrtval;
  const double St = 8. * G * sqrtval;
  const double kn = 4. / 3. * Y * sqrtval;
  const double kt = St;
  const double s56 = 0.91287092917527685576161630466800355658790782499663875;
  const double gamman = -2. * s56 * Bt * sqrt(Sn * meff);
  const double gammat = d.tdamp ? -2. * s56 * Bt * sqrt(St * meff) : 0.0;
  const double Fn_damping = -gamman * vn;
  const double Fn_contact = kn * deltan;
  double Fn = Fn_damping + Fn_contact;
  if (d.limit && Fn < 0.0) Fn = 0.0;
  const double Fn_i[3] = {Fn * en[0], Fn * en[1], Fn * en[2]};
  double ti[3];
  cross(xci, Fn_i, ti);
  for (int k = 0; k < 3; k++) {
    dFi[k] = Fn * en[k];
    dTi[k] = ti[k];
  }
  if (!is_wall) {
    const double Fn_j[3] = {-Fn_i[0], -Fn_i[1], -Fn_i[2]};
    cross(xcj, Fn_j, dTj);
  }
  if (shearupdate) {
    shear[0] += vt[0] * d.dt;
    shear[1] += vt[1] * d.dt;
    shear[2] += vt[2] * d.dt;
    const double rsht = shear[0] * en[0] + shear[1] * en[1] + shear[2] * en[2];
    shear[0] -= rsht * en[0];
    shear[1] -= rsht * en[1];
    shear[2] -= rsht * en[2];
  }
  const double shrmag = sqrt(shear[0] * shear[0] + shear[1] * shear[1] + shear[2] * shear[2]);
  const double xmu = d.mu[tt];
  double Ft[3] = {-(kt * shear[0]), -(kt * shear[1]), -(kt * shear[2])};
  const double Ft_shear = kt * shrmag;
  const double Ft_friction = xmu * fabs(Fn);
  if (Ft_shear > Ft_friction) {
    if (shrmag != 0.0) {
      const double ratio = Ft_friction / Ft_shear;
      Ft[0] *= ratio;
      Ft[1] *= ratio;
      Ft[2] *= ratio;
      if (shearupdate) {
        shear[0] = -Ft[0] / kt;
        shear[1] = -Ft[1] / kt;
        shear[2] = -Ft[2] / kt;
      }
    } else {
      Ft[0] = Ft[1] = Ft[2] = 0.0;
    }
  } else {
    Ft[0] -= gammat * vt[0];
    Ft[1] -= gammat * vt[1];
    Ft[2] -= gammat * vt[2];
  }
  double tti[3];
  cross(xci, Ft, tti);
  for (int k = 0; k < 3; k++) {
    dFi[k] += Ft[k];
    dTi[k] += tti[k];
  }
  if (!is_wall) {
    const double Ftj[3] = {-Ft[0], -Ft[1], -Ft[2]};
    double ttj[3];
    cross(xcj, Ftj, ttj);
    for (int k = 0; k < 3; k++) dTj[k] += ttj[k];
  }
}

// PairGran compute loop body (pair_gran_base.h:255-430) for one pair with the
// superquadric surface model: checkSurfaceIntersect
// (surface_model_superquadric.h:108-205) on overlapping bounding spheres, the
// force chain on intersecting surfaces, surfacesClose (zero shear) otherwise
__global__ __launch_bounds__(kBlk) void k_pairs(Dev d, int shearupdate) {
  const int64_t p = (int64_t)blockIdx.x * kBlk + threadIdx.x;
  if (p >= d.np) return;
  double *rec = d.rec + (size_t)kRec * p;
  double out[kRec];
  for (int k = 0; k < kRec; k++) out[k] = 0.0;
  const int i = d.pi[p], j = d.pj[p];
  double xj[3], del[3];
  image_of(d, i, j, xj, del);
  const double rsq = del[0] * del[0] + del[1] * del[1] + del[2] * del[2];
  const double radsum = d.rb[i] + d.rb[j];
  if (rsq < radsum * radsum) {
    double *h = d.hist + (size_t)kHist * p;
    double poseA[12], poseB[12];
    load_pose(d, i, d.x + 3 * i, poseA);
    load_pose(d, j, xj, poseB);
    sqd::Sq A, B;
    sqd::init(A, poseA);
    sqd::init(B, poseB);
    const int state = (int)(h[0] + 0.5);
    int start = (int)(h[1] + 0.5);
    const int obb = state == 2 ? 1 : sqd::obb_intersect(A, B, start);
    int contact = 0, iters = 0, fail = 0, fallback = 0;
    double cp[3] = {h[2], h[3], h[4]}, fi = 0.0, fj = 0.0, en[3] = {0.0, 0.0, 0.0}, dn = 0.0;
    if (obb) {
      const double ri = cbrt(A.shape[0] * A.shape[1] * A.shape[2]);
      const double rj = cbrt(B.shape[0] * B.shape[1] * B.shape[2]);
      const double ratio = ri / (ri + rj);
      int fresh = state == 0;
      const int slot = (fresh && d.init) ? d.fresh_idx[p] : -1;
      if (!fresh || slot >= 0) {
        double est[3];
        if (fresh) {
          for (int c = 0; c < 3; c++) est[c] = d.init[3 * (size_t)slot + c];
        } else {
          const double mi = d.mass[i], mj = d.mass[j];
          const double *wi = d.w + 3 * i, *wj = d.w + 3 * j;
          const double xci[3] = {cp[0] - A.center[0], cp[1] - A.center[1], cp[2] - A.center[2]};
          const double xcj[3] = {cp[0] - B.center[0], cp[1] - B.center[1], cp[2] - B.center[2]};
          double ri_[3], rj_[3];
          cross(wi, xci, ri_);
          cross(wj, xcj, rj_);
          for (int c = 0; c < 3; c++) {
            const double vi = d.v[3 * i + c] + ri_[c], vj = d.v[3 * j + c] + rj_[c];
            const double veff = (vi * mi + vj * mj) / (mi + mj);
            est[c] = cp[c] + veff * d.dt;
          }
        }
        iters = sqd::contact_point(A, B, ratio, est, cp, fi, fj, fail, d.max_iter);
        if (fresh) {
          const double merit = sqd::merit_of(A.gradient, B.gradient, fi, fj);
          if (fail || !(merit < 1e-10)) fallback = 1;
          fresh = fallback;
        }
      }
      if (fresh) iters += sqd::contact_point_fresh(A, B, cp, fi, fj, fail, d.max_iter);
      if (state == 0) {
        atomicAdd(&d.stats[kStatFresh], 1ull);
        atomicAdd(&d.stats[kStatFreshIters], (unsigned long long)iters);
        if (fallback) atomicAdd(&d.stats[kStatFallback], 1ull);
      }
      contact = (fi > fj ? fi : fj) < 0.0;
      if (contact) {
        for (int c = 0; c < 3; c++) en[c] = B.gradient[c] - A.gradient[c];
        const double nrm = sqrt(en[0] * en[0] + en[1] * en[1] + en[2] * en[2]);
        const double inv = nrm == 0. ? 0. : 1. / nrm;
        for (int c = 0; c < 3; c++) en[c] *= inv;
        const double men[3] = {-en[0], -en[1], -en[2]};
        double cpa[3], cpb[3];
        h[5] = sqd::line_intersection(A, cp, men, h[5], cpa);
        h[6] = sqd::line_intersection(B, cp, en, h[6], cpb);
        dn = fabs(h[5]) + fabs(h[6]);
        h[0] = 2.0;
      } else {
        h[0] = 1.0;
      }
      h[2] = cp[0];
      h[3] = cp[1];
      h[4] = cp[2];
    } else {
      h[0] = 0.0;
    }
    h[1] = (double)start;
    atomicAdd(&d.stats[kStatChecked], 1ull);
    if (iters) atomicAdd(&d.stats[kStatNewton], (unsigned long long)iters);
    if (contact) {
      d.flag[p] = 1;
      atomicAdd(&d.stats[kStatPairTouch], 1ull);
      const double rvi = d.rv[i], rvj = d.rv[j];
      const double mi = d.mass[i], mj = d.mass[j];
      const double xci[3] = {cp[0] - d.x[3 * i], cp[1] - d.x[3 * i + 1], cp[2] - d.x[3 * i + 2]};
      const double xcj[3] = {cp[0] - xj[0], cp[1] - xj[1], cp[2] - xj[2]};
      contact_force(d, d.type[i], d.type[j], 0, rvi * rvj / (rvi + rvj), mi * mj / (mi + mj), dn, en,
                    d.v + 3 * i, d.v + 3 * j, d.w + 3 * i, d.w + 3 * j, xci, xcj, shearupdate, h + 7, out,
                    out + 3, out + 6);
    } else {
      h[7] = h[8] = h[9] = 0.0;
    }
  }
  for (int k = 0; k < kRec; k++) rec[k] = out[k];
}

// Superquadric::plane_intersection (superquadric.cpp:440-507): the point of
// maximal penetration; returns 1 when particle and plane intersect
__device__ __forceinline__ int plane_intersection(const sqd::Sq &s, const double *ng, const double *x0g,
                                                  double *pmax) {
  double nl[3], x0l[3], pm[3], pl[3];
  tmv(s.rot, ng, nl);
  const double d0[3] = {x0g[0] - s.center[0], x0g[1] - s.center[1], x0g[2] - s.center[2]};
  tmv(s.rot, d0, x0l);
  if (nl[0] * x0l[0] + nl[1] * x0l[1] + nl[2] * x0l[2] < 0.0) {
    nl[0] = -nl[0];
    nl[1] = -nl[1];
    nl[2] = -nl[2];
  }
  const double n1 = s.blk[0], n2 = s.blk[1];
  const double nx = nl[0], ny = nl[1], nz = nl[2];
  const double a = s.shape[0], b = s.shape[1], c = s.shape[2];
  const double dd = nl[0] * x0l[0] + nl[1] * x0l[1] + nl[2] * x0l[2];
  if (sqd::comp_double(nx, 0.0, 1e-10) && sqd::comp_double(ny, 0.0, 1e-10)) {
    pm[0] = pm[1] = 0.0;
    pm[2] = 1.0;
    pl[0] = pl[1] = 0.0;
    pl[2] = dd / nz;
  } else {
    if (fabs(nx) > fabs(ny)) {
      const double alpha = sqd::pow_abs((ny * b) / (nx * a), 1.0 / (n2 - 1.0));
      const double gamma1 = 1.0 + sqd::pow_abs(alpha, n2);
      const double gamma = sqd::pow_abs(gamma1, n1 / n2 - 1.0);
      const double beta = sqd::pow_abs((nz * c) / (nx * a) * gamma, 1.0 / (n1 - 1.0));
      pm[0] = 1.0 / sqd::pow_abs(sqd::pow_abs(gamma1, n1 / n2) + sqd::pow_abs(beta, n1), 1.0 / n1);
      pm[1] = alpha * pm[0];
      pm[2] = beta * pm[0];
      pl[0] = fabs(dd) / (fabs(nx) * a + alpha * fabs(ny) * b + beta * fabs(nz) * c);
      pl[1] = alpha * pl[0];
      pl[2] = beta * pl[0];
    } else {
      const double alpha = sqd::pow_abs((nx * a) / (ny * b), 1.0 / (n2 - 1.0));
      const double gamma1 = 1.0 + sqd::pow_abs(alpha, n2);
      const double gamma = sqd::pow_abs(gamma1, n1 / n2 - 1.0);
      const double beta = sqd::pow_abs((nz * c) / (ny * b) * gamma, 1.0 / (n1 - 1.0));
      pm[1] = 1.0 / sqd::pow_abs(sqd::pow_abs(gamma1, n1 / n2) + sqd::pow_abs(beta, n1), 1.0 / n1);
      pm[0] = alpha * pm[1];
      pm[2] = beta * pm[1];
      pl[1] = fabs(dd) / (fabs(ny) * b + alpha * fabs(nx) * a + beta * fabs(nz) * c);
      pl[0] = alpha * pl[1];
      pl[2] = beta * pl[1];
    }
    for (int k = 0; k < 3; k++) pl[k] *= s.shape[k] * sqd::sgn(nl[k]);
  }
  for (int k = 0; k < 3; k++) pm[k] *= s.shape[k] * sqd::sgn(nl[k]);
  const double ff = sqd::shape_local(s, pl);
  double g[3];
  mv(s.rot, pm, g);
  for (int k = 0; k < 3; k++) pmax[k] = g[k] + s.center[k];
  return ff < 0.0;
}

// FixWallGran::post_force_primitive for a superquadric (fix_wall_gran.cpp:
// 1024-1153) and the wall variant of the force chain (fix_wall_gran_base.h:148-361)
__device__ __forceinline__ void wall_one(const Dev &d, int i, int k, int shearupdate, double *F, double *T) {
  double *sh = d.whist + 3 * ((size_t)i * d.nplanes + k);
  const int ax = d.pax[k];
  const double pos = d.ppos[k];
  double xi[3];
  for (int m = 0; m < 3; m++) xi[m] = d.x[3 * i + m];
  double delta[3] = {0.0, 0.0, 0.0}, deltan;
  delta[ax] = pos - xi[ax];
  if (xi[ax] > pos)
    deltan = xi[ax] - pos - d.rb[i];
  else
    deltan = pos - xi[ax] - d.rb[i];
  if (deltan > 0.0) {
    sh[0] = sh[1] = sh[2] = 0.0;
    return;
  }
  const double scp[3] = {xi[0] + delta[0], xi[1] + delta[1], xi[2] + delta[2]};
  double pose[12];
  load_pose(d, i, xi, pose);
  sqd::Sq S;
  sqd::init(S, pose);
  double cp[3];
  const int hit = plane_intersection(S, delta, scp, cp);
  if (!hit) {
    sh[0] = sh[1] = sh[2] = 0.0;
    return;
  }
  // point_wall_projection (math_extra_liggghts_nonspherical.cpp:48-66)
  const double A = delta[0], Bc = delta[1], C = delta[2];
  const double dist = fabs((A * (cp[0] - scp[0]) + Bc * (cp[1] - scp[1]) + C * (cp[2] - scp[2])) /
                           sqrt(A * A + Bc * Bc + C * C));
  atomicAdd(&d.stats[kStatWallTouch], 1ull);
  const double sd[3] = {-delta[0], -delta[1], -delta[2]};
  const double dinv = 1.0 / sqrt(sd[0] * sd[0] + sd[1] * sd[1] + sd[2] * sd[2]);
  const double en[3] = {sd[0] * dinv, sd[1] * dinv, sd[2] * dinv};
  const double xci[3] = {cp[0] - xi[0], cp[1] - xi[1], cp[2] - xi[2]};
  const double vwall[3] = {0.0, 0.0, 0.0};
  double dF[3], dT[3], unused[3];
  double shl[3] = {sh[0], sh[1], sh[2]};
  contact_force(d, d.type[i], d.ptype[k], 1, d.rv[i], d.mass[i], dist, en, d.v + 3 * i, vwall, d.w + 3 * i,
                nullptr, xci, nullptr, shearupdate, shl, dF, dT, unused);
  sh[0] = shl[0];
  sh[1] = shl[1];
  sh[2] = shl[2];
  for (int m = 0; m < 3; m++) {
    F[m] += dF[m];
    T[m] += dT[m];
  }
}

// force_clear, the pair records of particle i (as j, then as i, by pair
// index: the order the reference's loop over i adds them), post_force (walls,
// gravity in script order), and FixNVEAsphereBase::final_integrate,
// integration_scheme 1 (fix_nve_asphere_base.cpp:529-610)
__global__ __launch_bounds__(kBlk) void k_gather(Dev d, int shearupdate, int integrate) {
  const int i = blockIdx.x * kBlk + threadIdx.x;
  if (i >= d.n) return;
  double F[3] = {0.0, 0.0, 0.0}, T[3] = {0.0, 0.0, 0.0};
  for (int r = d.rfirst[i]; r < d.rfirst[i + 1]; r++) {
    const double *rec = d.rec + (size_t)kRec * d.rlist[r];
    for (int m = 0; m < 3; m++) {
      F[m] += -rec[m];
      T[m] += rec[6 + m];
    }
  }
  for (int p = d.first[i]; p < d.first[i + 1]; p++) {
    const double *rec = d.rec + (size_t)kRec * p;
    for (int m = 0; m < 3; m++) {
      F[m] += rec[m];
      T[m] += rec[3 + m];
    }
  }
  const double mi = d.mass[i];
  if (d.grav_first && d.has_grav)
    for (int m = 0; m < 3; m++) F[m] += mi * d.g[m];
  for (int k = 0; k < d.nplanes; k++) wall_one(d, i, k, shearupdate, F, T);
  if (!d.grav_first && d.has_grav)
    for (int m = 0; m < 3; m++) F[m] += mi * d.g[m];
  for (int m = 0; m < 3; m++) {
    d.f[3 * i + m] = F[m];
    d.t[3 * i + m] = T[m];
  }
  if (!integrate) return;
  const double dtf = 0.5 * d.dt, dtf2 = 2.0 * dtf;
  double q[4], L[3], I[3], R[9], mbody[3], conjqm[4], tbody[3], fquat[4], wbody[3], w[3];
  for (int m = 0; m < 4; m++) q[m] = d.q[4 * i + m];
  for (int m = 0; m < 3; m++) {
    L[m] = d.L[3 * i + m];
    I[m] = d.I[3 * i + m];
  }
  const double dtfm = dtf / mi;
  rot_of(q, R);
  for (int m = 0; m < 3; m++) d.v[3 * i + m] += dtfm * F[m];
  tmv(R, T, tbody);
  tmv(R, L, mbody);
  quatvec(q, mbody, conjqm);
  for (int m = 0; m < 4; m++) conjqm[m] *= 2.0;
  quatvec(q, tbody, fquat);
  for (int m = 0; m < 4; m++) conjqm[m] += dtf2 * fquat[m];
  invquatvec(q, conjqm, mbody);
  for (int m = 0; m < 3; m++) mbody[m] *= 0.5;
  for (int m = 0; m < 3; m++) wbody[m] = mbody[m] / I[m];
  mv(R, mbody, L);
  mv(R, wbody, w);
  for (int m = 0; m < 3; m++) {
    d.L[3 * i + m] = L[m];
    d.w[3 * i + m] = w[m];
  }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
hipError_t launch_initial(const Dev &d, int *flag, hipStream_t s) {
  if (d.n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_initial, dim3(nblk(d.n, kBlk)), dim3(kBlk), 0, s, d, flag);
  return hipGetLastError();
}

hipError_t launch_wrap(const Dev &d, hipStream_t s) {
  if (d.n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_wrap, dim3(nblk(d.n, kBlk)), dim3(kBlk), 0, s, d);
  return hipGetLastError();
}

hipError_t launch_bin(const Dev &d, const Grid &g, hipStream_t s) {
  hipLaunchKernelGGL(k_zero_int, dim3(nblk(g.ncell + 1, kBlk)), dim3(kBlk), 0, s, g.count, g.ncell + 1);
  hipLaunchKernelGGL(k_zero_int, dim3(nblk(g.ncell + 1, kBlk)), dim3(kBlk), 0, s, g.fill, g.ncell + 1);
  hipLaunchKernelGGL(k_bin_count, dim3(nblk(d.n, kBlk)), dim3(kBlk), 0, s, d, g);
  hipError_t e = dem::launch_exclusive_scan(g.count, g.start, g.scan_tmp, g.ncell, s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_bin_fill, dim3(nblk(d.n, kBlk)), dim3(kBlk), 0, s, d, g);
  hipLaunchKernelGGL(k_bin_sort, dim3(nblk(g.ncell, kBlk)), dim3(kBlk), 0, s, g);
  return hipGetLastError();
}

hipError_t launch_pair_count(const Dev &d, const Grid &g, const Build &b, hipStream_t s) {
  hipLaunchKernelGGL(k_pair_count, dim3(nblk(d.n, kBlk)), dim3(kBlk), 0, s, d, g, b);
  return hipGetLastError();
}

hipError_t launch_pair_fill(const Dev &d, const Grid &g, const Build &b, hipStream_t s) {
  hipLaunchKernelGGL(k_pair_fill, dim3(nblk(d.n, kBlk)), dim3(kBlk), 0, s, d, g, b);
  return hipGetLastError();
}

hipError_t launch_reverse(const Dev &d, const Build &b, int *scan_tmp, hipStream_t s) {
  const int np = (int)d.np;
  hipLaunchKernelGGL(k_zero_int, dim3(nblk(d.n + 1, kBlk)), dim3(kBlk), 0, s, b.rcount, d.n + 1);
  hipLaunchKernelGGL(k_zero_int, dim3(nblk(d.n + 1, kBlk)), dim3(kBlk), 0, s, b.rfill, d.n + 1);
  if (np > 0) hipLaunchKernelGGL(k_rev_count, dim3(nblk(np, kBlk)), dim3(kBlk), 0, s, b, np);
  hipError_t e = dem::launch_exclusive_scan(b.rcount, b.rfirst, scan_tmp, d.n, s);
  if (e != hipSuccess) return e;
  if (np > 0) hipLaunchKernelGGL(k_rev_fill, dim3(nblk(np, kBlk)), dim3(kBlk), 0, s, b, np);
  hipLaunchKernelGGL(k_rev_sort, dim3(nblk(d.n, kBlk)), dim3(kBlk), 0, s, b, d.n);
  return hipGetLastError();
}

hipError_t launch_zero_stats(unsigned long long *stats, hipStream_t s) {
  hipLaunchKernelGGL(k_zero_stats, dim3(1), dim3(64), 0, s, stats);
  return hipGetLastError();
}

hipError_t launch_accumulate(const unsigned long long *stats, unsigned long long *totals, hipStream_t s) {
  hipLaunchKernelGGL(k_accumulate, dim3(1), dim3(64), 0, s, stats, totals);
  return hipGetLastError();
}

hipError_t launch_fresh_scan(const Dev &d, int *fresh_idx, int *fresh_count, double *pose_a, double *pose_b,
                             hipStream_t s) {
  hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(kBlk), 0, s, fresh_count, 1);
  if (d.np <= 0) return hipGetLastError();
  hipLaunchKernelGGL(k_fresh_scan, dim3(nblk(d.np, kBlk)), dim3(kBlk), 0, s, d, fresh_idx, fresh_count, pose_a,
                     pose_b);
  return hipGetLastError();
}

hipError_t launch_pairs(const Dev &d, int shearupdate, hipStream_t s) {
  if (d.np <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_pairs, dim3(nblk(d.np, kBlk)), dim3(kBlk), 0, s, d, shearupdate);
  return hipGetLastError();
}

hipError_t launch_gather(const Dev &d, int shearupdate, int integrate, hipStream_t s) {
  if (d.n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_gather, dim3(nblk(d.n, kBlk)), dim3(kBlk), 0, s, d, shearupdate, integrate);
  return hipGetLastError();
}

}  // namespace sqdk
