// HIP kernels of the DEM step for gfx950 (MI355X).
//
// Layout: particles live in Morton-cell order (re-sorted at every neighbor
// rebuild), as packed double4 rows (pos+radius, vel+mass, omega+type) so a
// neighbor gather is one or two 32-byte loads. The neighbor list is a full
// list in column-major ELL form (entry k of particle i at k*cap+i), so the 64
// lanes of a wavefront read consecutive words; bit 31 of an entry is the
// contact flag. Each direction of a pair keeps its own copy of the tangential
// history, in the same ELL layout; the two copies are exact mirror images
// (x_i - x_j is formed so that swapping i and j negates it bit for bit), so the
// list needs no atomics and every particle's force is summed by one thread.
//
// The file is compiled with -ffp-contract=off: in FP64 mode every product and
// sum rounds as in the CPU reference, so one pair's contribution is bitwise
// equal to the reference's; only the order of the per-particle sums differs.
#include "dem_device.h"
#include "mesh_device.h"

namespace dem {

constexpr double kPi = 3.14159265358979323846;  // M_PI

// ---------------------------------------------------------------------------
// helpers
// ---------------------------------------------------------------------------

// Workgroups are dispatched round-robin over the 8 XCDs (block b -> XCD b%8).
// Remap so that each XCD walks one contiguous range of Morton-ordered
// particles: neighbors of a particle then sit in the same XCD's L2.
__device__ __forceinline__ int xcd_block(int nblocks) {
  const int b = blockIdx.x;
  const int xcd = b & 7, slot = b >> 3;
  const int q = nblocks >> 3, r = nblocks & 7;
  return xcd * q + (xcd < r ? xcd : r) + slot;
}

__device__ __forceinline__ int cell_coord(const DevParams &p, int d, double x) {
  int c = (int)floor((x - p.lo[d]) * p.cinv[d]);
  c = c < 0 ? 0 : c;
  c = c >= p.nc[d] ? p.nc[d] - 1 : c;
  return c;
}

// The minimum-image shift of a difference x_i - x_j: a component beyond half
// the period of a periodic dimension moves by one period. Exactly
// antisymmetric under i <-> j.
__device__ __forceinline__ void apply_image(const DevParams &p, double &dx, double &dy,
                                            double &dz) {
  if (p.per[0]) {
    if (dx > p.halfL[0]) dx -= p.L[0];
    else if (dx < -p.halfL[0]) dx += p.L[0];
  }
  if (p.per[1]) {
    if (dy > p.halfL[1]) dy -= p.L[1];
    else if (dy < -p.halfL[1]) dy += p.L[1];
  }
  if (p.per[2]) {
    if (dz > p.halfL[2]) dz -= p.L[2];
    else if (dz < -p.halfL[2]) dz += p.L[2];
  }
}

// Whether apply_image would change the difference. The callers apply the
// shift only on lanes that need it, so a wavefront with no periodic-image
// partner skips that code; apply_image is the identity elsewhere, so the
// results are the same bits either way.
__device__ __forceinline__ bool needs_image(const DevParams &p, double dx, double dy, double dz) {
  return (p.per[0] & (fabs(dx) > p.halfL[0])) | (p.per[1] & (fabs(dy) > p.halfL[1])) |
         (p.per[2] & (fabs(dz) > p.halfL[2]));
}

// x_i - x_j with the minimum-image shift
__device__ __forceinline__ void min_image(const DevParams &p, double xi, double yi, double zi,
                                          double xj, double yj, double zj, double &dx,
                                          double &dy, double &dz) {
  dx = xi - xj;
  dy = yi - yj;
  dz = zi - zj;
  apply_image(p, dx, dy, dz);
}


template <typename T> struct OmgT;
template <> struct OmgT<double> { typedef double4 type; };
template <> struct OmgT<float> { typedef float4 type; };

// three history values of a record (a 4-vector load / store; the fourth
// value of a record is padding and is written as zero)
template <typename T>
__device__ __forceinline__ void load_rec3(const T *p, T v[3]) {
  typedef typename OmgT<T>::type V4;
  const V4 r = *(const V4 *)p;
  v[0] = r.x;
  v[1] = r.y;
  v[2] = r.z;
}
template <typename T>
__device__ __forceinline__ void store_rec3(T *p, const T v[3]) {
  typedef typename OmgT<T>::type V4;
  V4 r;
  r.x = v[0];
  r.y = v[1];
  r.z = v[2];
  r.w = T(0);
  *(V4 *)p = r;
}

// ---------------------------------------------------------------------------
// contact law, in the order of ContactModel::surfacesIntersect
// (contact_models.h:228-238): SurfaceModel<DEFAULT> (surface_model_default.h:
// 140-205) -> NormalModel<HERTZ> (normal_model_hertz.h:202-405) ->
// TangentialModel<HISTORY> (tangential_model_history.h:130-356) ->
// RollingModel<CDT> (rolling_model_cdt.h:89-167). deltan is formed by the
// caller in FP64. With GEN the normal model may instead be HOOKE
// (normal_model_hooke.h:227-436), HERTZ_STIFFNESS
// (normal_model_hertz_stiffness.h:106-202) or HOOKE_STIFFNESS
// (normal_model_hooke_stiffness.h:194-365), CohesionModel<SJKR>
// (cohesion_model_sjkr.h:86-124) may follow it, and the tangential model may
// be NO_HISTORY (tangential_model_no_history.h:112-242).
// ---------------------------------------------------------------------------
template <typename T>
struct CIn {
  T radi, radj, rinv, meff, deltan, deltan_in;
  T en[3], del[3];
  T vr[3];
  T wi[3], wj[3];
  T Y, G, beta, mu, mur;
  // GEN only: sibling-model coefficients, characteristic velocity, and the
  // sjkr contact-circle area (formed in FP64 by the caller)
  T kn, kt, gn, gt, ced, loge, cv, acont;
};

// contact-model switches (the kModel* bits)
struct Model {
  int rolling, tdamp, limitf, torsion;
  int normal, nohist, sjkr, kt2kn, absdamp, treduce, epsd2;
};

__device__ __forceinline__ Model model_from_bits(int b) {
  return Model{b & 1, (b >> 1) & 1, (b >> 2) & 1, (b >> 3) & 1,
               (b >> kModelNormalShift) & 3, (b & kModelNoHistory) != 0, (b & kModelSjkr) != 0,
               (b & kModelKtToKn) != 0, (b & kModelAbsDamping) != 0, (b & kModelTangReduce) != 0,
               (b & kModelEpsd2) != 0};
}

__device__ __forceinline__ double tpow(double x, double y) { return pow(x, y); }
__device__ __forceinline__ float tpow(float x, float y) { return powf(x, y); }

// sibling-model coefficients of type pair mt (GEN contacts only)
template <typename T>
__device__ __forceinline__ void load_model_coeffs(const MatTables &mt_, const DevParams &p,
                                                  const Model &m, const int mt, CIn<T> &c) {
  c.cv = (T)p.char_vel;
  c.kn = c.kt = c.gn = c.gt = c.ced = c.loge = T(0);
  if (m.normal == kNormalHooke) c.loge = sizeof(T) == 8 ? (T)mt_.loge[mt] : (T)mt_.floge[mt];
  if (m.normal >= kNormalHertzStiffness) {
    if (sizeof(T) == 8) {
      c.kn = (T)mt_.kn[mt];
      c.kt = (T)mt_.kt[mt];
      c.gn = (T)mt_.gn[mt];
      c.gt = (T)mt_.gt[mt];
    } else {
      c.kn = (T)mt_.fkn[mt];
      c.kt = (T)mt_.fkt[mt];
      c.gn = (T)mt_.fgn[mt];
      c.gt = (T)mt_.fgt[mt];
    }
  }
  if (m.sjkr) c.ced = sizeof(T) == 8 ? (T)mt_.ced[mt] : (T)mt_.fced[mt];
}

template <typename T, bool WALL, bool GEN = false>
__device__ __forceinline__ void contact_chain(const Model &p, const CIn<T> &c, T dt,
                                              int shearupdate, T sh[3], T rh[3], T dF[3],
                                              T dTi[3], T dTj[3]) {
  const T enx = c.en[0], eny = c.en[1], enz = c.en[2];
  const T vr1 = c.vr[0], vr2 = c.vr[1], vr3 = c.vr[2];
  const T vn = vr1 * enx + vr2 * eny + vr3 * enz;
  const T vn1 = vn * enx, vn2 = vn * eny, vn3 = vn * enz;
  const T vt1 = vr1 - vn1, vt2 = vr2 - vn2, vt3 = vr3 - vn3;
  const T deltan = c.deltan;
  const T dx = c.del[0], dy = c.del[1], dz = c.del[2];
  const T rinv = c.rinv;
  T wr1, wr2, wr3, cri, crj = T(0);
  if (WALL) {
    const T cr = c.radi - T(0.5) * c.deltan_in;
    wr1 = cr * c.wi[0] * rinv;
    wr2 = cr * c.wi[1] * rinv;
    wr3 = cr * c.wi[2] * rinv;
    cri = cr;
  } else {
    cri = c.radi - T(0.5) * deltan;
    crj = c.radj - T(0.5) * deltan;
    wr1 = (cri * c.wi[0] + crj * c.wj[0]) * rinv;
    wr2 = (cri * c.wi[1] + crj * c.wj[1]) * rinv;
    wr3 = (cri * c.wi[2] + crj * c.wj[2]) * rinv;
  }
  const T vtr1 = vt1 - (dz * wr2 - dy * wr3);
  const T vtr2 = vt2 - (dx * wr3 - dz * wr1);
  const T vtr3 = vt3 - (dy * wr1 - dx * wr2);

  const T reff = WALL ? c.radi : (c.radi * c.radj / (c.radi + c.radj));
  T kn, kt, gamman, gammat;
  if (GEN && p.normal == kNormalHooke) {
    // viscous off
    const T sqrtval = sqrt(reff);
    kn = T(16.) / T(15.) * sqrtval * c.Y *
         tpow(T(15.) * c.meff * c.cv * c.cv / (T(16.) * sqrtval * c.Y), T(0.2));
    kt = kn;
    if (p.kt2kn) kt *= T(0.285714286);
    const T lsq = c.loge * c.loge;
    const T pi = T(3.14159265358979323846);
    gamman = sqrt(T(4.) * c.meff * kn * lsq / (lsq + pi * pi));
    gammat = p.tdamp ? gamman : T(0);
  } else if (GEN && p.normal == kNormalHertzStiffness) {
    const T polyhertz = sqrt(reff * deltan);
    kn = polyhertz * c.kn;
    kt = polyhertz * c.kt;
    gamman = polyhertz * c.meff * c.gn;
    gammat = p.tdamp ? polyhertz * c.meff * c.gt : T(0);
  } else if (GEN && p.normal == kNormalHookeStiffness) {
    kn = c.kn;
    kt = c.kt;
    if (p.absdamp) {
      gamman = c.gn;
      gammat = c.gt;
    } else {
      gamman = c.meff * c.gn;
      gammat = c.meff * c.gt;
    }
    if (!p.tdamp) gammat = T(0);
  } else {
    // Hertz
    const T sqrtval = sqrt(reff * deltan);
    const T Sn = T(2.) * c.Y * sqrtval;
    const T St = T(8.) * c.G * sqrtval;
    kn = T(4.) / T(3.) * c.Y * sqrtval;
    kt = St;
    const T s56 = T(0.91287092917527685576161630466800355658790782499663875);
    gamman = T(-2.) * s56 * c.beta * sqrt(Sn * c.meff);
    gammat = p.tdamp ? T(-2.) * s56 * c.beta * sqrt(St * c.meff) : T(0);
  }
  const T Fn_damping = -gamman * vn;
  const T Fn_contact = kn * deltan;
  T Fn = Fn_damping + Fn_contact;
  if (p.limitf && Fn < T(0)) Fn = T(0);
  dF[0] = Fn * enx;
  dF[1] = Fn * eny;
  dF[2] = Fn * enz;

  // cohesion: energy density times the contact-circle area, attractive
  if (GEN && p.sjkr) {
    const T Fn_coh = -c.ced * c.acont;
    if (p.treduce) Fn += Fn_coh;  // the tangential model sees the reduced force
    dF[0] += Fn_coh * enx;
    dF[1] += Fn_coh * eny;
    dF[2] += Fn_coh * enz;
  }

  T Ft1, Ft2, Ft3;
  if (GEN && p.nohist) {
    // viscous friction gammat |vtr| capped by Coulomb; nothing is stored
    const T vrel = sqrt(vtr1 * vtr1 + vtr2 * vtr2 + vtr3 * vtr3);
    const T Ft_friction = c.mu * fabs(Fn);
    T gamma = T(0);
    if (Ft_friction < gammat * vrel)
      gamma = Ft_friction / vrel;
    else
      gamma = gammat;
    Ft1 = -gamma * vtr1;
    Ft2 = -gamma * vtr2;
    Ft3 = -gamma * vtr3;
  } else {
  // tangential history
  if (shearupdate) {
    sh[0] += vtr1 * dt;
    sh[1] += vtr2 * dt;
    sh[2] += vtr3 * dt;
    const T rsht = sh[0] * enx + sh[1] * eny + sh[2] * enz;
    sh[0] -= rsht * enx;
    sh[1] -= rsht * eny;
    sh[2] -= rsht * enz;
  }
  const T shrmag = sqrt(sh[0] * sh[0] + sh[1] * sh[1] + sh[2] * sh[2]);
  Ft1 = -(kt * sh[0]);
  Ft2 = -(kt * sh[1]);
  Ft3 = -(kt * sh[2]);
  const T Ft_shear = kt * shrmag;
  const T Ft_friction = c.mu * fabs(Fn);
  if (Ft_shear > Ft_friction) {
    if (shrmag != T(0)) {
      const T ratio = Ft_friction / Ft_shear;
      Ft1 *= ratio;
      Ft2 *= ratio;
      Ft3 *= ratio;
      if (shearupdate) {
        sh[0] = -Ft1 / kt;
        sh[1] = -Ft2 / kt;
        sh[2] = -Ft3 / kt;
      }
    } else {
      Ft1 = Ft2 = Ft3 = T(0);
    }
  } else {
    Ft1 -= (gammat * vtr1);
    Ft2 -= (gammat * vtr2);
    Ft3 -= (gammat * vtr3);
  }
  }
  const T tor1 = eny * Ft3 - enz * Ft2;
  const T tor2 = enz * Ft1 - enx * Ft3;
  const T tor3 = enx * Ft2 - eny * Ft1;
  dF[0] += Ft1;
  dF[1] += Ft2;
  dF[2] += Ft3;
  dTi[0] = T(0) + -cri * tor1;
  dTi[1] = T(0) + -cri * tor2;
  dTi[2] = T(0) + -cri * tor3;
  if (!WALL) {
    dTj[0] = T(0) + -crj * tor1;
    dTj[1] = T(0) + -crj * tor2;
    dTj[2] = T(0) + -crj * tor3;
  }

  if (GEN && p.epsd2) {
    // RollingModel<EPSD2> (rolling_model_epsd2.h:98-221): elastic spring
    // kr = kt R*^2 on the relative spin (its twist about en dropped unless
    // torsionTorque), capped at mu_r |Fn| R*; rh carries the torque
    T w0, w1, w2;
    if (WALL) {
      w0 = wr1;
      w1 = wr2;
      w2 = wr3;
    } else {
      w0 = c.wi[0] - c.wj[0];
      w1 = c.wi[1] - c.wj[1];
      w2 = c.wi[2] - c.wj[2];
    }
    if (!p.torsion) {
      const T d = w0 * enx + w1 * eny + w2 * enz;
      const T n0 = enx * d, n1 = eny * d, n2 = enz * d;
      w0 = w0 - n0;
      w1 = w1 - n1;
      w2 = w2 - n2;
    }
    const T kr = kt * reff * reff;
    const T dk = dt * kr;
    T rt0 = rh[0] + w0 * dk, rt1 = rh[1] + w1 * dk, rt2 = rh[2] + w2 * dk;
    const T rmag = sqrt(rt0 * rt0 + rt1 * rt1 + rt2 * rt2);
    const T rmax = fabs(Fn) * reff * c.mur;
    if (rmag > rmax) {
      const T factor = rmax / rmag;
      rt0 *= factor;
      rt1 *= factor;
      rt2 *= factor;
    }
    if (shearupdate) {
      rh[0] = rt0;
      rh[1] = rt1;
      rh[2] = rt2;
    }
    dTi[0] -= rt0;
    dTi[1] -= rt1;
    dTi[2] -= rt2;
    if (!WALL) {
      dTj[0] += rt0;
      dTj[1] += rt1;
      dTj[2] += rt2;
    }
  } else if (p.rolling) {
    // rolling friction, constant directional torque
    const T rmu = c.mur;
    T rt0 = T(0), rt1 = T(0), rt2 = T(0);
    if (WALL) {
      const T wrmag = sqrt(wr1 * wr1 + wr2 * wr2 + wr3 * wr3);
      if (wrmag > T(0)) {
        const T Fnw = deltan * kn;
        rt0 = rmu * Fnw * wr1 / wrmag * reff;
        rt1 = rmu * Fnw * wr2 / wrmag * reff;
        rt2 = rmu * Fnw * wr3 / wrmag * reff;
      }
    } else {
      const T w0 = c.wi[0] - c.wj[0], w1 = c.wi[1] - c.wj[1], w2 = c.wi[2] - c.wj[2];
      const T mag = sqrt(w0 * w0 + w1 * w1 + w2 * w2);
      if (mag > T(0)) {
        const T fac = rmu * kn * deltan * reff / mag;
        rt0 = w0 * fac;
        rt1 = w1 * fac;
        rt2 = w2 * fac;
      }
    }
    if (!p.torsion) {
      const T d = rt0 * enx + rt1 * eny + rt2 * enz;
      const T n0 = enx * d, n1 = eny * d, n2 = enz * d;
      rt0 = rt0 - n0;
      rt1 = rt1 - n1;
      rt2 = rt2 - n2;
    }
    dTi[0] -= rt0;
    dTi[1] -= rt1;
    dTi[2] -= rt2;
    if (!WALL) {
      dTj[0] += rt0;
      dTj[1] += rt1;
      dTj[2] += rt2;
    }
  }
}

// ---------------------------------------------------------------------------
// mesh walls for one particle: FixWallGran::post_force_mesh (fix_wall_gran.cpp:
// 845-1018) with FixContactHistoryMesh::markAllContacts / handleContact /
// cleanUpContacts (fix_contact_history_mesh_I.h:51-200,
// fix_contact_history_mesh.cpp:405-464). Candidates are visited in ascending
// triangle index, mesh by mesh, as the reference's triangle loop visits them.
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool tri_coplanar(const MeshStep &M, int a, int b) {
  for (int k = M.copl_off[a]; k < M.copl_off[a + 1]; k++)
    if (M.copl_idx[k] == b) return true;
  return false;
}

template <typename T, bool GEN, typename W>
__device__ __forceinline__ void mesh_contacts(const StepArgs &a, int i, const double4 &pi,
                                     const double4 &vi, const W &wi, int irow, T &fx, T &fy,
                                     T &fz, T &tx, T &ty, T &tz,
                                     unsigned long long &nwall) {
  const DevParams &p = a.prm;
  const MeshStep &M = a.mesh;
  const int nc = M.ncand[i];
  if (nc == 0) return;
  const size_t cap = (size_t)a.cap;
  const double px[3] = {pi.x, pi.y, pi.z};
  int c = 0;
  for (int m = 0; m < p.nmesh; m++) {
    int *slot = a.aux.mslot + (size_t)m * kMeshSlots * cap + i;
    // a slot's history: the shear, then (hist_w 6) the epsd2 rolling torque
    const size_t mhw = GEN ? (size_t)p.hist_w : 3;
    T *hist = (T *)a.aux.mhist + (size_t)m * kMeshSlots * mhw * cap + i;
    const int tend = p.mesh_tri0[m] + p.mesh_ntri[m];
    const Model mdl = model_from_bits(p.mesh_model[m]);
    const bool roll = GEN && mdl.epsd2;
    const int mt = irow + p.mesh_wtype[m];
    unsigned keep = 0u;
    for (; c < nc; c++) {
      const int t = M.cand[(size_t)c * cap + i];
      if (t >= tend) break;
      const TriGeom &tg = M.tri[t];
      Vec3 delta;
      Bary bary;
      int bs;
      const double deltan = tri_sphere(tg, p.mesh_precision, pi.w, px, delta, bary, bs);
      if (deltan > p.cutneighmax) continue;
      if (!(deltan <= 0 || deltan < (p.cdf - 1.0) * pi.w)) continue;
      // handleContact
      int s = -1;
      for (int k = 0; k < kMeshSlots; k++)
        if (slot[(size_t)k * cap] == t) {
          s = k;
          break;
        }
      bool fresh = false;
      if (s >= 0) {
        keep |= 1u << s;
      } else {
        bool skip = false;
        if (bs == 7)
          for (int k = 0; k < kMeshSlots; k++) {
            const int st = slot[(size_t)k * cap];
            if (st >= 0 && st != t && ((keep >> k) & 1u) && tri_coplanar(M, st, t)) skip = true;
          }
        if (skip) continue;
        for (int k = 0; k < kMeshSlots; k++)
          if (slot[(size_t)k * cap] == -1) {
            s = k;
            break;
          }
        if (s < 0) {
          atomicMax(&a.ctrl->slot_overflow, 1);
          continue;
        }
        slot[(size_t)s * cap] = t;
        keep |= 1u << s;
        fresh = true;
      }
      T sh[3], rh[3] = {T(0), T(0), T(0)};
      if (fresh) {
        sh[0] = sh[1] = sh[2] = T(0);
        if (bs == 7)  // checkCoplanarContactHistory: the last coplanar slot wins
          for (int k = 0; k < kMeshSlots; k++) {
            const int st = slot[(size_t)k * cap];
            if (st >= 0 && st != t && tri_coplanar(M, st, t))
              for (int d = 0; d < 3; d++) {
                sh[d] = hist[(size_t)(k * mhw + d) * cap];
                if (roll) rh[d] = hist[(size_t)(k * mhw + 3 + d) * cap];
              }
          }
      } else {
        for (int d = 0; d < 3; d++) {
          sh[d] = hist[(size_t)(s * mhw + d) * cap];
          if (roll) rh[d] = hist[(size_t)(s * mhw + 3 + d) * cap];
        }
      }
      // within (cdf - 1) r of the triangle but not touching: compute_force with
      // intersectflag false, whose surfacesClose zeroes the slot's shear
      // (fix_wall_gran_base.h:283-289, tangential_model_history.h:358-369) and
      // the epsd2 rolling torque (rolling_model_epsd2.h:144-151)
      const bool band = !(deltan <= 0);
      if (band && !mdl.nohist) sh[0] = sh[1] = sh[2] = T(0);
      if (band && roll) rh[0] = rh[1] = rh[2] = T(0);
      double vw[3] = {0., 0., 0.};
      if (p.mesh_moving[m])
        for (int k = 0; k < 3; k++)
          vw[k] = (bary.b0 * tg.vn[k] + bary.b1 * tg.vn[3 + k] + bary.b2 * tg.vn[6 + k]);
      if (deltan <= 0) {
        // sidata.deltan = -deltan, sidata.delta = -delta; Walls::Granular::compute_force
        const double sdn = -deltan;
        const double r = pi.w - sdn;
        const double rinv = 1.0 / r;
        const double sd0 = -delta.x, sd1 = -delta.y, sd2 = -delta.z;
        CIn<T> cc;
        cc.radi = (T)pi.w;
        cc.radj = T(0);
        cc.rinv = (T)rinv;
        cc.meff = (T)vi.w;
        cc.deltan = (T)(pi.w - r);
        cc.deltan_in = (T)sdn;
        cc.del[0] = (T)sd0;
        cc.del[1] = (T)sd1;
        cc.del[2] = (T)sd2;
        cc.en[0] = (T)(sd0 * rinv);
        cc.en[1] = (T)(sd1 * rinv);
        cc.en[2] = (T)(sd2 * rinv);
        cc.vr[0] = (T)(vi.x - vw[0]);
        cc.vr[1] = (T)(vi.y - vw[1]);
        cc.vr[2] = (T)(vi.z - vw[2]);
        cc.wi[0] = (T)wi.x;
        cc.wi[1] = (T)wi.y;
        cc.wi[2] = (T)wi.z;
        cc.wj[0] = cc.wj[1] = cc.wj[2] = T(0);
        if (sizeof(T) == 8) {
          cc.Y = (T)a.mat.Yeff[mt];
          cc.G = (T)a.mat.Geff[mt];
          cc.beta = (T)a.mat.beta[mt];
          cc.mu = (T)a.mat.mu[mt];
          cc.mur = (T)a.mat.mur[mt];
        } else {
          cc.Y = (T)a.mat.fYeff[mt];
          cc.G = (T)a.mat.fGeff[mt];
          cc.beta = (T)a.mat.fbeta[mt];
          cc.mu = (T)a.mat.fmu[mt];
          cc.mur = (T)a.mat.fmur[mt];
        }
        if (GEN) {
          load_model_coeffs(a.mat, p, mdl, mt, cc);
          // sphere-wall contact area (cohesion_model_sjkr.h:95-96)
          cc.acont = (T)((pi.w * pi.w - r * r) * kPi);
        }
        T dF[3], dTi[3], dTj[3];
        contact_chain<T, true, GEN>(mdl, cc, (T)p.dt, a.shearupdate, sh, rh, dF, dTi, dTj);
        if (a.count && ((p.mesh_stress >> m) & 1)) {
          // MeshModuleStress::add_particle_contribution (mesh_module_stress.cpp:
          // 276-300): the force on the mesh is minus the particle's, its torque
          // is taken about the reference point at the contact point x + delta
          const double fm[3] = {-(double)dF[0], -(double)dF[1], -(double)dF[2]};
          const double rp[3] = {px[0] + delta.x - p.mesh_pref[m][0], px[1] + delta.y - p.mesh_pref[m][1],
                                px[2] + delta.z - p.mesh_pref[m][2]};
          double *mf = a.ctrl->mesh_force[m];
          atomicAdd(&mf[0], fm[0]);
          atomicAdd(&mf[1], fm[1]);
          atomicAdd(&mf[2], fm[2]);
          atomicAdd(&mf[3], rp[1] * fm[2] - rp[2] * fm[1]);
          atomicAdd(&mf[4], rp[2] * fm[0] - rp[0] * fm[2]);
          atomicAdd(&mf[5], rp[0] * fm[1] - rp[1] * fm[0]);
        }
        fx += dF[0];
        fy += dF[1];
        fz += dF[2];
        tx += dTi[0];
        ty += dTi[1];
        tz += dTi[2];
        nwall++;
      }
      if (a.shearupdate || fresh || band)
        for (int d = 0; d < 3; d++) {
          hist[(size_t)(s * mhw + d) * cap] = sh[d];
          if (mhw == 6) hist[(size_t)(s * mhw + 3 + d) * cap] = rh[d];
        }
    }
    // cleanUpContacts: slots not touched this pass are freed
    for (int k = 0; k < kMeshSlots; k++)
      if (!((keep >> k) & 1u) && slot[(size_t)k * cap] != -1) slot[(size_t)k * cap] = -1;
  }
}

// ---------------------------------------------------------------------------
// fix move/mesh rotate: MeshMoverRotate::initial_integrate
// (mesh_mover_rotation.cpp:110-139), MultiNodeMesh::rotate
// (multi_node_mesh_I.h:620-660) and MultiNodeMesh::decideRebuild
// (multi_node_mesh_I.h:792-825)
// ---------------------------------------------------------------------------
// translating movers (MeshMoverLinear / MeshMoverWiggle / MeshMoverVibLin::
// initial_integrate, mesh_mover_linear.cpp:88-101, 203-218, 275-305): nodes
// and centre shifted by v dt (TrackingMesh::move -> MultiNodeMesh::move,
// multi_node_mesh_I.h:502-527; edge vectors and normals are translation
// invariant), node velocities reset to v (FixMoveMesh::initial_integrate
// zeroes them first). Wiggle: v = A (w cos(w t)); viblin: v = axis vA, the
// factors of run step `step` from the motion table.
__device__ __forceinline__ void mesh_translate_tri(const MeshMoveArgs &a, const TriGeom &src,
                                                   TriGeom &dst, const int m, const long long step) {
  double vN[3];
  if (a.kind[m] == kMoverLinear) {
    for (int k = 0; k < 3; k++) vN[k] = a.tvec[m][k];
  } else {
    const double sc = a.motion[(step * kMaxMeshes + m) * kMotionStride + 7];
    for (int k = 0; k < 3; k++) vN[k] = a.tvec[m][k] * sc;
  }
  double dx[3];
  for (int k = 0; k < 3; k++) dx[k] = vN[k] * a.dt;
  if (a.dst) {  // fields a translation leaves unchanged
    for (int k = 0; k < 9; k++) dst.ev[k] = src.ev[k];
    for (int k = 0; k < 3; k++) dst.sn[k] = src.sn[k];
    for (int k = 0; k < 3; k++) dst.el[k] = src.el[k];
    dst.rb = src.rb;
    dst.obtuse = src.obtuse;
    dst.flags = src.flags;
    dst.mesh = src.mesh;
    dst.pad = src.pad;
  }
  for (int j = 0; j < 3; j++)
    for (int k = 0; k < 3; k++) dst.n[3 * j + k] = src.n[3 * j + k] + dx[k];
  for (int k = 0; k < 3; k++) dst.c[k] = src.c[k] + dx[k];
  for (int j = 0; j < 3; j++)
    for (int k = 0; k < 3; k++) dst.vn[3 * j + k] = 0.0 + vN[k];
}

// moves triangle t of a moving mesh one step (to its geometry of run step
// `step`); returns whether one of its nodes is now more than skin/2 from where
// it was at the last rebuild
__device__ __forceinline__ bool mesh_move_tri(const MeshMoveArgs &a, const int t,
                                              const long long step) {
  const TriGeom &src = a.tri[t];
  const int m = src.mesh;
  if (!((a.moving_mask >> m) & 1)) return false;
  TriGeom &dst = a.dst ? a.dst[t] : a.tri[t];
  if (a.kind[m] != kMoverRotate && a.kind[m] != kMoverVibRot) {
    mesh_translate_tri(a, src, dst, m, step);
    bool moved = false;
    if (a.decide)
      for (int j = 0; j < 3; j++) {
        const double dx = dst.n[3 * j] - a.node_last[9 * t + 3 * j];
        const double dy = dst.n[3 * j + 1] - a.node_last[9 * t + 3 * j + 1];
        const double dz = dst.n[3 * j + 2] - a.node_last[9 * t + 3 * j + 2];
        if (dx * dx + dy * dy + dz * dz > a.triggersq) moved = true;
      }
    return moved;
  }
  // rotate: constant increment; vibrot (MeshMoverVibRot::initial_integrate,
  // mesh_mover_rotation.cpp:397-430): increment vR dt and angular velocity
  // axis vR of this step from the motion table
  const bool vib = a.kind[m] == kMoverVibRot;
  const double *mot = vib ? a.motion + (step * kMaxMeshes + m) * kMotionStride : nullptr;
  const double *dQ = vib ? mot : a.dQ[m];
  const double *omegaVec = vib ? mot + 4 : a.omegaVec[m];
  const double *o = a.origin[m];
  double c[3] = {0., 0., 0.};
  double nd[9];
  for (int j = 0; j < 3; j++) {
    double v[3] = {src.n[3 * j], src.n[3 * j + 1], src.n[3 * j + 2]};
    if (a.trans[m]) d3_sub(v, o, v);
    vec_quat_rotate(v, dQ);
    if (a.trans[m])
      for (int k = 0; k < 3; k++) v[k] = v[k] + o[k];
    for (int k = 0; k < 3; k++) {
      nd[3 * j + k] = v[k];
      c[k] = v[k] + c[k];
    }
  }
  d3_div(c, 3.0);
  double ev[9], sn[3] = {src.sn[0], src.sn[1], src.sn[2]};
  for (int j = 0; j < 3; j++) {
    double v[3] = {src.ev[3 * j], src.ev[3 * j + 1], src.ev[3 * j + 2]};
    vec_quat_rotate(v, dQ);
    for (int k = 0; k < 3; k++) ev[3 * j + k] = v[k];
  }
  vec_quat_rotate(sn, dQ);
  if (a.dst) {  // fields the rotation leaves unchanged
    for (int k = 0; k < 3; k++) dst.el[k] = src.el[k];
    dst.rb = src.rb;
    dst.obtuse = src.obtuse;
    dst.flags = src.flags;
    dst.mesh = src.mesh;
    dst.pad = src.pad;
  }
  for (int k = 0; k < 9; k++) dst.n[k] = nd[k];
  for (int k = 0; k < 3; k++) dst.c[k] = c[k];
  for (int k = 0; k < 9; k++) dst.ev[k] = ev[k];
  for (int k = 0; k < 3; k++) dst.sn[k] = sn[k];
  for (int j = 0; j < 3; j++) {
    double rPA[3], vRot[3];
    d3_sub(&nd[3 * j], o, rPA);
    d3_cross(omegaVec, rPA, vRot);
    for (int k = 0; k < 3; k++) dst.vn[3 * j + k] = 0.0 + vRot[k];
  }
  bool moved = false;
  if (a.decide)
    for (int j = 0; j < 3; j++) {
      const double dx = nd[3 * j] - a.node_last[9 * t + 3 * j];
      const double dy = nd[3 * j + 1] - a.node_last[9 * t + 3 * j + 1];
      const double dz = nd[3 * j + 2] - a.node_last[9 * t + 3 * j + 2];
      if (dx * dx + dy * dy + dz * dz > a.triggersq) moved = true;
    }
  return moved;
}

// step: the step the moved geometry belongs to
__device__ __forceinline__ void mesh_move_block(const MeshMoveArgs &a, const int t,
                                                const long long step) {
  const bool moved = t < a.ntri && mesh_move_tri(a, t, step);
  if (__ballot(moved) != 0ull && (threadIdx.x & 63) == 0)
    atomicMin((unsigned long long *)&a.ctrl->stop_at, (unsigned long long)(step + 1));
}

// The surfacesClose branch of the pair loop (pair_gran_base.h:414-417) for the
// entries of bit mask `close` (group k0, entries ent): within cdf * radsum but
// not touching. ContactModel::surfacesClose zeroes the shear history
// (tangential_model_history.h:358-369) and the epsd2 rolling torque
// (rolling_model_epsd2.h:144-151) and leaves the contact flag nonzero only if
// its normal-model bit is set: hertz, hooke and hooke/stiffness set it here
// (normal_model_hertz.h:407-411 and likewise), hertz/stiffness keeps it only
// from a carry-over at the last build.
template <typename T, bool GEN, int GROUP>
__device__ __forceinline__ void close_entries(const StepArgs &a, const Model &pm, const int i,
                                              const int k0, uint32_t close,
                                              const uint32_t (&ent)[GROUP], T *hist,
                                              const size_t cap, const size_t hw) {
  const bool keeps = !(GEN && pm.nohist);
  const bool normal_bit = pm.normal != kNormalHertzStiffness;
  while (close) {
    const int u = __builtin_ctz(close);
    close &= close - 1u;
    const int k = k0 + u;
    const uint32_t e = ent[u];
    if (keeps) {
      const T z[3] = {T(0), T(0), T(0)};
      const size_t hb = hist_at(k, i, cap, hist_rec((int)hw));
      store_rec3(hist + hb, z);
      if (hw == 6) store_rec3(hist + hb + 4, z);
    }
    const uint32_t ne = (normal_bit || (e & kCarryBit)) ? (e | kFlagBit) : (e & ~kFlagBit);
    if (ne != e) a.list.nbr[(size_t)k * cap + i] = ne;
  }
}

// One touching list entry (k, e) of particle i: pair_gran_base.h:356-430 with
// the contact chain, the history read / write of the entry and its flag.
// Returns the contribution to i's force and torque.
// STORE = false: the evaluation of compute pair/gran/local, which writes
// neither history nor flag.
// pre: the entry's shear history when the caller has loaded it already.
template <typename T, bool GEN, typename W, bool STORE = true>
__device__ __forceinline__ void pair_contact(const StepArgs &a, const Model &pm, const int i,
                                             const int k, const uint32_t e, const double4 &pi,
                                             const double4 &vi, const W &wi, const int irow,
                                             T dF[3], T dTi[3], const T *pre = nullptr) {
  typedef typename OmgT<T>::type omg_t;
  const DevParams &p = a.prm;
  const size_t cap = (size_t)a.cap;
  const size_t hw = GEN ? (size_t)p.hist_w : 3;  // history values per contact
  T *hist = (T *)a.list.hist;
  const int j = (int)(e & kIndexMask);
  const double4 pj = a.in.pos[j];
  const double4 vj = a.in.vel[j];
  const omg_t wj = ((const omg_t *)a.in.omg)[j];
  double dx, dy, dz;
  min_image(p, pi.x, pi.y, pi.z, pj.x, pj.y, pj.z, dx, dy, dz);
  const double rsq = dx * dx + dy * dy + dz * dz;
  const double radsum = pi.w + pj.w;
  const int jtype = (int)wj.w & kTypeMask;
  const double r = sqrt(rsq);
  const double rinv = 1.0 / r;
  CIn<T> c;
  c.radi = (T)pi.w;
  c.radj = (T)pj.w;
  c.rinv = (T)rinv;
  c.meff = (T)(vi.w * vj.w / (vi.w + vj.w));
  c.deltan = (T)(radsum - r);
  c.deltan_in = T(0);
  c.en[0] = (T)(dx * rinv);
  c.en[1] = (T)(dy * rinv);
  c.en[2] = (T)(dz * rinv);
  c.del[0] = (T)dx;
  c.del[1] = (T)dy;
  c.del[2] = (T)dz;
  c.vr[0] = (T)(vi.x - vj.x);
  c.vr[1] = (T)(vi.y - vj.y);
  c.vr[2] = (T)(vi.z - vj.z);
  c.wi[0] = (T)wi.x;
  c.wi[1] = (T)wi.y;
  c.wi[2] = (T)wi.z;
  c.wj[0] = (T)wj.x;
  c.wj[1] = (T)wj.y;
  c.wj[2] = (T)wj.z;
  const int mt = irow + jtype;
  if (sizeof(T) == 8) {
    c.Y = (T)a.mat.Yeff[mt];
    c.G = (T)a.mat.Geff[mt];
    c.beta = (T)a.mat.beta[mt];
    c.mu = (T)a.mat.mu[mt];
    c.mur = (T)a.mat.mur[mt];
  } else {
    c.Y = (T)a.mat.fYeff[mt];
    c.G = (T)a.mat.fGeff[mt];
    c.beta = (T)a.mat.fbeta[mt];
    c.mu = (T)a.mat.fmu[mt];
    c.mur = (T)a.mat.fmur[mt];
  }
  if (GEN) {
    load_model_coeffs(a.mat, p, pm, mt, c);
    // sphere-sphere contact-circle area (cohesion_model_sjkr.h:97-98)
    const double ri = pi.w, rj = pj.w;
    c.acont = (T)(-kPi / 4 * ((r - ri - rj) * (r + ri - rj) * (r - ri + rj) * (r + ri + rj)) / (r * r));
  }
  const bool flagged = (e & kFlagBit) != 0;
  // tangential no_history keeps no history: its slots hold the zeros
  // written at the first touch and are neither read nor rewritten
  const bool keeps = !(GEN && pm.nohist);
  T sh[3], rh[3] = {T(0), T(0), T(0)};
  const size_t hb = hist_at(k, i, cap, hist_rec((int)hw));
  if (flagged && keeps) {
    if (pre) {
      sh[0] = pre[0];
      sh[1] = pre[1];
      sh[2] = pre[2];
    } else {
      load_rec3(hist + hb, sh);
    }
    if (GEN && pm.epsd2) load_rec3(hist + hb + 4, rh);
  } else {
    sh[0] = sh[1] = sh[2] = T(0);
  }
  T dTj[3];
  contact_chain<T, false, GEN>(pm, c, (T)p.dt, STORE ? a.shearupdate : 0, sh, rh, dF, dTi, dTj);
  if (!STORE) return;
  if ((a.shearupdate && keeps) || !flagged) {
    store_rec3(hist + hb, sh);
    if (GEN && hw == 6) store_rec3(hist + hb + 4, rh);
  }
  if (!flagged) a.list.nbr[(size_t)k * cap + i] = e | kFlagBit;
}

// Broad phase of list entries k0 .. k0+GROUP-1 of particle i (cnt entries):
// the index loads of the group are independent, then its position gathers,
// so a wavefront has GROUP memory requests in flight instead of one
// dependent chain per entry. Entries past cnt gather particle `pad` (a valid
// index: i itself, or 0 for a lane past the last particle) and are ignored.
// Returns the touching entries as bits; entries in the surfacesClose band are
// handled here.
// With hpre (Hertz / history kernels), the shear histories of the group's
// entries flagged in contact at the last step are loaded beside the position
// gathers: a history record is read once per step, so it comes from HBM, and
// loading it here keeps that latency out of the contact law's chain.
template <typename T, bool GEN, int GROUP>
__device__ __forceinline__ uint32_t broad_phase(const StepArgs &a, const Model &pm, const int i,
                                                const int pad, const int cnt, const int k0,
                                                const double4 &pi, uint32_t (&ent)[GROUP],
                                                T (*hpre)[3] = nullptr) {
  const DevParams &p = a.prm;
  const size_t cap = (size_t)a.cap;
#pragma unroll
  for (int u = 0; u < GROUP; u++)
    ent[u] = (k0 + u < cnt) ? a.list.nbr[(size_t)(k0 + u) * cap + i] : (uint32_t)pad;
  double4 pj[GROUP];
#pragma unroll
  for (int u = 0; u < GROUP; u++) pj[u] = a.in.pos[ent[u] & kIndexMask];
  if (hpre) {
    const T *hist = (const T *)a.list.hist;
#pragma unroll
    for (int u = 0; u < GROUP; u++) {
      hpre[u][0] = hpre[u][1] = hpre[u][2] = T(0);
      if (k0 + u < cnt && (ent[u] & kFlagBit)) load_rec3(hist + hist_at(k0 + u, i, cap, 4), hpre[u]);
    }
  }
  double dx[GROUP], dy[GROUP], dz[GROUP];
  bool wrap = false;
#pragma unroll
  for (int u = 0; u < GROUP; u++) {
    dx[u] = pi.x - pj[u].x;
    dy[u] = pi.y - pj[u].y;
    dz[u] = pi.z - pj[u].z;
    wrap |= needs_image(p, dx[u], dy[u], dz[u]);
  }
  // the minimum-image shift only on lanes with a periodic-image partner
  if (wrap) {
#pragma unroll
    for (int u = 0; u < GROUP; u++) apply_image(p, dx[u], dy[u], dz[u]);
  }
  uint32_t bits = 0u, close = 0u;
#pragma unroll
  for (int u = 0; u < GROUP; u++) {
    const double rsq = dx[u] * dx[u] + dy[u] * dy[u] + dz[u] * dz[u];
    const double radsum = pi.w + pj[u].w;
    if (k0 + u < cnt) {
      if (rsq < radsum * radsum)
        bits |= 1u << u;
      else if (p.close_band && rsq < p.cdm * radsum * radsum)
        close |= 1u << u;
    }
  }
  if (close) {
    const size_t hw = GEN ? (size_t)p.hist_w : 3;
    close_entries<T, GEN, GROUP>(a, pm, i, k0, close, ent, (T *)a.list.hist, cap, hw);
  }
  return bits;
}

// Orders LDS accesses of one wavefront: the writes before it are visible to
// the reads after it, by any lane of the same wavefront.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// LDS of one wavefront for the compacted pair phase: the touching entries of
// a group of the 64 particles (neighbor entry, owner lane << 16 | list slot)
// and the contributions of one pass of the contact law.
template <typename T, int GROUP>
struct WaveLds {
  uint2 queue[64 * GROUP];
  T contrib[6][64];
};

// Pair forces of the wavefront's 64 particles with the touching entries
// compacted: each group of GROUP list entries per particle is broad-phased by
// its own lane, the touching entries of all lanes are ranked by ballot prefix
// sums into an LDS queue slot-major (entry u of every lane, in lane order,
// before entry u+1 of any lane), and the whole wavefront runs the contact law
// on 64 queued entries at a time. Slot-major order puts the entries of
// neighboring particles at the same list slot side by side, so the history
// and owner accesses of one pass fall in a few cache lines. Each particle then
// adds its entries' contributions in list order (its queue positions rise with
// the slot), so the sums are bitwise those of the per-lane path. Every lane of
// the wavefront takes part; lanes past the last particle have cnt = 0.
template <typename T, bool GEN, int GROUP>
__device__ __forceinline__ void pair_forces_wave(const StepArgs &a, const Model &pm, const int i,
                                                 const int pad, const int cnt, const double4 &pi,
                                                 T &fx, T &fy,
                                                 T &fz, T &tx, T &ty, T &tz,
                                                 unsigned long long &my_contacts,
                                                 WaveLds<T, GROUP> &L) {
  typedef typename OmgT<T>::type omg_t;
  const DevParams &p = a.prm;
  const int lane = threadIdx.x & 63;
  const int ibase = i - lane;  // particle of lane 0
  const unsigned long long below = (1ull << lane) - 1ull;
  int kmax = cnt;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const int o = __shfl_xor(kmax, off, 64);
    kmax = o > kmax ? o : kmax;
  }
  for (int k0 = 0; k0 < kmax; k0 += GROUP) {
    uint32_t ent[GROUP];
    const uint32_t bits = broad_phase<T, GEN, GROUP>(a, pm, i, pad, cnt, k0, pi, ent);
    // queue position of this lane's entry u: the touching entries of slots
    // below u, then those of lower lanes at slot u
    int pos[GROUP];
    int total = 0;
#pragma unroll
    for (int u = 0; u < GROUP; u++) {
      const unsigned long long m = __ballot((bits >> u) & 1u);
      pos[u] = total + __popcll(m & below);
      total += __popcll(m);
    }
#pragma unroll
    for (int u = 0; u < GROUP; u++)
      if ((bits >> u) & 1u) L.queue[pos[u]] = make_uint2(ent[u], ((uint32_t)lane << 16) | (uint32_t)(k0 + u));
    my_contacts += (unsigned long long)__popc(bits);
    wave_sync();
    for (int c0 = 0; c0 < total; c0 += 64) {
      const int q = c0 + lane;
      T dF[3] = {T(0), T(0), T(0)}, dTi[3] = {T(0), T(0), T(0)};
      if (q < total) {
        const uint2 qe = L.queue[q];
        const int io = ibase + (int)(qe.y >> 16);
        const int k = (int)(qe.y & 0xffffu);
        const double4 po = a.in.pos[io];
        const double4 vo = a.in.vel[io];
        const omg_t wo = ((const omg_t *)a.in.omg)[io];
        const int wb = (int)wo.w;
        const int orow = (wb >> kTypeBits) * p.mat_stride + (wb & kTypeMask) * p.ntp1;
        pair_contact<T, GEN>(a, pm, io, k, qe.x, po, vo, wo, orow, dF, dTi);
      }
      L.contrib[0][lane] = dF[0];
      L.contrib[1][lane] = dF[1];
      L.contrib[2][lane] = dF[2];
      L.contrib[3][lane] = dTi[0];
      L.contrib[4][lane] = dTi[1];
      L.contrib[5][lane] = dTi[2];
      wave_sync();
      // this pass's entries of this particle, in list order
#pragma unroll
      for (int u = 0; u < GROUP; u++) {
        const int l = pos[u] - c0;
        if (((bits >> u) & 1u) && l >= 0 && l < 64) {
          fx += L.contrib[0][l];
          fy += L.contrib[1][l];
          fz += L.contrib[2][l];
          tx += L.contrib[3][l];
          ty += L.contrib[4][l];
          tz += L.contrib[5][l];
        }
      }
      wave_sync();
    }
  }
}

// LDS of one wavefront for the row-queued pair phase: the records of the
// wavefront's 64 particles, the first kRowStash touching list entries of each
// particle in the current round of list slots, the queue of the round's
// touching entries particle by particle (owner lane << 21 | touching rank << 16
// | list slot), and the contributions of one pass of the contact law.
constexpr int kRowQueue = 256;
constexpr int kRowStash = 8;
template <typename T>
struct RowLds {
  double4 pos[64];
  double4 vel[64];
  typename OmgT<T>::type omg[64];
  uint32_t stash[kRowStash][64];
  uint32_t queue[kRowQueue];
  T contrib[6][64];
};

template <typename T, int GROUP, int WAVE>
struct PairLdsT {
  typedef WaveLds<T, GROUP> type;
};
template <typename T, int GROUP>
struct PairLdsT<T, GROUP, 2> {
  typedef RowLds<T> type;
};

// Pair forces of the wavefront's 64 particles with the touching entries of
// whole rows queued. Each lane broad-phases rounds of 32 list slots of its
// particle (GROUP entries in flight at a time), keeps the touching ones as
// bits and stashes their list entries in LDS; an exclusive scan of the
// counts gives each particle a contiguous range of the queue, its entries in
// slot order; the wavefront then runs the contact law on 64 queued entries at
// a time, with the owners' records and list entries read from LDS, and each
// particle adds the contributions of its range in that order, so its sums are
// bitwise those of the per-lane path. A round of touching entries longer than
// the queue is taken in windows of kRowQueue. Every lane of the wavefront
// takes part; lanes past the last particle have cnt = 0. The particle records
// are in L.pos / L.vel / L.omg, staged by the caller.
template <typename T, bool GEN, int GROUP>
__device__ __forceinline__ void pair_forces_rowq(const StepArgs &a, const Model &pm, const int i,
                                                 const int pad, const int cnt, T &fx, T &fy,
                                                 T &fz, T &tx, T &ty, T &tz,
                                                 unsigned long long &my_contacts, RowLds<T> &L) {
  static_assert(32 % GROUP == 0, "a round holds whole groups");
  typedef typename OmgT<T>::type omg_t;
  const DevParams &p = a.prm;
  const size_t cap = (size_t)a.cap;
  const int lane = threadIdx.x & 63;
  const int ibase = i - lane;  // particle of lane 0
  int kmax = cnt;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const int o = __shfl_xor(kmax, off, 64);
    kmax = o > kmax ? o : kmax;
  }
  for (int r0 = 0; r0 < kmax; r0 += 32) {
    const int rend = kmax < r0 + 32 ? kmax : r0 + 32;
    uint32_t tmask = 0u;
    int c = 0;
    for (int k0 = r0; k0 < rend; k0 += GROUP) {
      uint32_t ent[GROUP];
      const double4 pi = L.pos[lane];
      uint32_t bits = broad_phase<T, GEN, GROUP>(a, pm, i, pad, cnt, k0, pi, ent);
      tmask |= bits << (k0 - r0);
      while (bits) {
        const int u = __builtin_ctz(bits);
        bits &= bits - 1u;
        uint32_t e = ent[0];
#pragma unroll
        for (int v = 1; v < GROUP; v++) e = u == v ? ent[v] : e;
        if (c < kRowStash) L.stash[c][lane] = e;
        c++;
      }
    }
    my_contacts += (unsigned long long)c;
    int incl = c;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const int o = __shfl_up(incl, off, 64);
      if (lane >= off) incl += o;
    }
    const int total = __shfl(incl, 63, 64);
    const int first = incl - c;
    for (int w0 = 0; w0 < total; w0 += kRowQueue) {
      const int wend = total < w0 + kRowQueue ? total : w0 + kRowQueue;
      uint32_t m = tmask;
      for (int q = first; m != 0u; q++) {
        const int u = __builtin_ctz(m);
        m &= m - 1u;
        if (q >= w0 && q < wend)
          L.queue[q - w0] = ((uint32_t)lane << 21) | ((uint32_t)(q - first) << 16) | (uint32_t)(r0 + u);
      }
      wave_sync();
      for (int c0 = w0; c0 < wend; c0 += 64) {
        const int q = c0 + lane;
        T dF[3] = {T(0), T(0), T(0)}, dTi[3] = {T(0), T(0), T(0)};
        if (q < wend) {
          const uint32_t qe = L.queue[q - w0];
          const int ol = (int)(qe >> 21);
          const int rank = (int)((qe >> 16) & 31u);
          const int io = ibase + ol;
          const int k = (int)(qe & 0xffffu);
          const uint32_t e = rank < kRowStash ? L.stash[rank][ol] : a.list.nbr[(size_t)k * cap + io];
          const double4 po = L.pos[ol];
          const double4 vo = L.vel[ol];
          const omg_t wo = L.omg[ol];
          const int wb = (int)wo.w;
          const int orow = (wb >> kTypeBits) * p.mat_stride + (wb & kTypeMask) * p.ntp1;
          pair_contact<T, GEN>(a, pm, io, k, e, po, vo, wo, orow, dF, dTi);
        }
        L.contrib[0][lane] = dF[0];
        L.contrib[1][lane] = dF[1];
        L.contrib[2][lane] = dF[2];
        L.contrib[3][lane] = dTi[0];
        L.contrib[4][lane] = dTi[1];
        L.contrib[5][lane] = dTi[2];
        wave_sync();
        // this particle's entries of the pass, in slot order
        const int lo = first > c0 ? first : c0;
        const int hi = first + c < c0 + 64 ? first + c : c0 + 64;
        for (int qq = lo; qq < hi; qq++) {
          const int l = qq - c0;
          fx += L.contrib[0][l];
          fy += L.contrib[1][l];
          fz += L.contrib[2][l];
          tx += L.contrib[3][l];
          ty += L.contrib[4][l];
          tz += L.contrib[5][l];
        }
        wave_sync();
      }
    }
  }
}

// One particle's step: contacts over its list, walls, meshes, gravity,
// integration and the displacement test. Accumulates its touching pair and
// wall contacts. WAVE 1: the pair phase is compacted over the wavefront
// group by group (pair_forces_wave), WAVE 2: row by row (pair_forces_rowq);
// WAVE 3: per lane as WAVE 0, with the shear histories of entries flagged in
// contact loaded in the broad phase (Hertz / history kernels);
// there every lane calls, active says whether it holds a particle.
// WAVE 0: each lane resolves its own touching entries.
template <typename T, bool MESH, int GROUP, bool GEN, int WAVE>
__device__ __forceinline__ void step_particle(const StepArgs &a, const int i, const bool active,
                                              unsigned long long &my_contacts,
                                              unsigned long long &my_wall, int &my_inst,
                                              bool &displaced, const long long ago_next,
                                              typename PairLdsT<T, GROUP, WAVE>::type *lds) {
  typedef typename OmgT<T>::type omg_t;
  const DevParams &p = a.prm;

  const int ii = active ? i : 0;
  double4 pi = a.in.pos[ii];
  double4 vi = a.in.vel[ii];
  omg_t wi = ((const omg_t *)a.in.omg)[ii];
  if constexpr (WAVE == 2) {
    // the records the row-queued pair phase reads; reloaded from LDS after it
    const int lane = threadIdx.x & 63;
    wave_sync();
    lds->pos[lane] = pi;
    lds->vel[lane] = vi;
    lds->omg[lane] = wi;
    wave_sync();
  }
  const int wbits = (int)wi.w;
  // material row of this particle: its instance's block, its type's row
  const int irow = (wbits >> kTypeBits) * p.mat_stride + (wbits & kTypeMask) * p.ntp1;
  my_inst = wbits >> kTypeBits;
  const int meta0 = active ? a.aux.meta[i] : 0;
  int meta = meta0;
  const int cnt = meta & kMetaCountMask;
  const T dt = (T)p.dt;
  const Model pm = model_from_bits(p.pair_model);
  T fx = T(0), fy = T(0), fz = T(0), tx = T(0), ty = T(0), tz = T(0);

  if (WAVE == 1 || WAVE == 2) {
    if constexpr (WAVE == 2) {
      pair_forces_rowq<T, GEN, GROUP>(a, pm, i, ii, cnt, fx, fy, fz, tx, ty, tz, my_contacts, *lds);
      const int lane = threadIdx.x & 63;
      pi = lds->pos[lane];
      vi = lds->vel[lane];
      wi = lds->omg[lane];
    } else
      pair_forces_wave<T, GEN, GROUP>(a, pm, i, ii, cnt, pi, fx, fy, fz, tx, ty, tz, my_contacts, *lds);
    if (!active) return;
  } else {
    // Touching entries of a group are resolved one by one, their list
    // entries selected from the group's registers.
    for (int k0 = 0; k0 < cnt; k0 += GROUP) {
      uint32_t ent[GROUP];
      constexpr bool PRE = WAVE == 3 && !GEN;
      T hpre[GROUP][3];
      uint32_t bits = broad_phase<T, GEN, GROUP>(a, pm, i, i, cnt, k0, pi, ent, PRE ? hpre : nullptr);
      while (bits) {
        const int u = __builtin_ctz(bits);
        const int k = k0 + u;
        bits &= bits - 1u;
        uint32_t e = ent[0];
        T sh[3] = {hpre[0][0], hpre[0][1], hpre[0][2]};
#pragma unroll
        for (int v = 1; v < GROUP; v++) {
          e = u == v ? ent[v] : e;
          if (PRE)
            for (int d = 0; d < 3; d++) sh[d] = u == v ? hpre[v][d] : sh[d];
        }
        T dF[3], dTi[3];
        pair_contact<T, GEN>(a, pm, i, k, e, pi, vi, wi, irow, dF, dTi, PRE ? sh : nullptr);
        fx += dF[0];
        fy += dF[1];
        fz += dF[2];
        tx += dTi[0];
        ty += dTi[1];
        tz += dTi[2];
        my_contacts++;
      }
    }
  }
  const size_t cap = (size_t)a.cap;
  const size_t hw = GEN ? (size_t)p.hist_w : 3;  // history values per contact

  // post_force fixes in script order: gravity (fix_gravity.cpp:331-338) and
  // plane walls (fix_wall_gran.cpp:1024-1153)
  for (int pass = 0; pass < 2; pass++) {
    const bool grav_pass = (pass == 0) == (p.grav_first != 0);
    if (grav_pass) {
      if (p.has_grav) {
        fx += (T)(vi.w * p.gacc[0]);
        fy += (T)(vi.w * p.gacc[1]);
        fz += (T)(vi.w * p.gacc[2]);
      }
      continue;
    }
    for (int w = 0; w < p.nwalls; w++) {
      const int ax = p.wall_axis[w];
      const double param = p.wall_pos[w];
      const double xa = ax == 0 ? pi.x : (ax == 1 ? pi.y : pi.z);
      double deltan, dlt;
      if (xa > param) {
        dlt = param - xa;
        deltan = xa - param - pi.w;
      } else {
        dlt = param - xa;
        deltan = param - xa - pi.w;
      }
      T *wh = (T *)a.aux.whist;
      const size_t wb = (size_t)w * hw * cap + i;
      const int bit = 1 << (kMetaWallShift + w);
      if (deltan > p.cutneighmax) continue;
      if (deltan <= 0) {
        // sidata.deltan = -deltan, sidata.delta = -delta; compute_force
        // (fix_wall_gran_base.h:179-188): r = radi - sidata.deltan
        const double sdn = -deltan;
        const double sd = -dlt;
        const double r = pi.w - sdn;
        const double rinv = 1.0 / r;
        CIn<T> c;
        c.radi = (T)pi.w;
        c.radj = T(0);
        c.rinv = (T)rinv;
        c.meff = (T)vi.w;
        c.deltan = (T)(pi.w - r);
        c.deltan_in = (T)sdn;
        // select instead of c.del[ax] = ...: a dynamic index would put the
        // struct in scratch memory
        const T sdT = (T)sd, enA = (T)(sd * rinv), enZ = (T)(0.0 * rinv);
        c.del[0] = ax == 0 ? sdT : T(0);
        c.del[1] = ax == 1 ? sdT : T(0);
        c.del[2] = ax == 2 ? sdT : T(0);
        c.en[0] = ax == 0 ? enA : enZ;
        c.en[1] = ax == 1 ? enA : enZ;
        c.en[2] = ax == 2 ? enA : enZ;
        c.vr[0] = (T)(vi.x - 0.0);
        c.vr[1] = (T)(vi.y - 0.0);
        c.vr[2] = (T)(vi.z - 0.0);
        c.wi[0] = (T)wi.x;
        c.wi[1] = (T)wi.y;
        c.wi[2] = (T)wi.z;
        c.wj[0] = c.wj[1] = c.wj[2] = T(0);
        const int mt = irow + p.wall_type[w];
        if (sizeof(T) == 8) {
          c.Y = (T)a.mat.Yeff[mt];
          c.G = (T)a.mat.Geff[mt];
          c.beta = (T)a.mat.beta[mt];
          c.mu = (T)a.mat.mu[mt];
          c.mur = (T)a.mat.mur[mt];
        } else {
          c.Y = (T)a.mat.fYeff[mt];
          c.G = (T)a.mat.fGeff[mt];
          c.beta = (T)a.mat.fbeta[mt];
          c.mu = (T)a.mat.fmu[mt];
          c.mur = (T)a.mat.fmur[mt];
        }
        const Model wm = model_from_bits(p.wall_model[w]);
        if (GEN) {
          load_model_coeffs(a.mat, p, wm, mt, c);
          c.acont = (T)((pi.w * pi.w - r * r) * kPi);  // cohesion_model_sjkr.h:95-96
        }
        const bool keeps = !(GEN && wm.nohist);  // as for pairs: no_history slots stay zero
        T sh[3], rh[3] = {T(0), T(0), T(0)};
        if ((meta & bit) && keeps) {
          sh[0] = wh[wb];
          sh[1] = wh[wb + cap];
          sh[2] = wh[wb + 2 * cap];
          if (GEN && wm.epsd2) {
            rh[0] = wh[wb + 3 * cap];
            rh[1] = wh[wb + 4 * cap];
            rh[2] = wh[wb + 5 * cap];
          }
        } else {
          sh[0] = sh[1] = sh[2] = T(0);
        }
        T dF[3], dTi[3], dTj[3];
        contact_chain<T, true, GEN>(wm, c, dt, a.shearupdate, sh, rh, dF, dTi, dTj);
        if (keeps || !(meta & bit)) {
          wh[wb] = sh[0];
          wh[wb + cap] = sh[1];
          wh[wb + 2 * cap] = sh[2];
          if (GEN && hw == 6) {
            wh[wb + 3 * cap] = rh[0];
            wh[wb + 4 * cap] = rh[1];
            wh[wb + 5 * cap] = rh[2];
          }
        }
        meta |= bit;
        fx += dF[0];
        fy += dF[1];
        fz += dF[2];
        tx += dTi[0];
        ty += dTi[1];
        tz += dTi[2];
        my_wall++;
      } else {
        meta &= ~bit;  // history zeroed when the contact ends (fix_wall_gran.cpp:1145-1151)
      }
    }
    if (MESH)
      mesh_contacts<T, GEN>(a, i, pi, vi, wi, irow, fx, fy, fz, tx, ty, tz, my_wall);
  }

  if (a.store_f) {
    a.fout[i] = (double)fx;
    a.fout[cap + i] = (double)fy;
    a.fout[2 * cap + i] = (double)fz;
    a.fout[3 * cap + i] = (double)tx;
    a.fout[4 * cap + i] = (double)ty;
    a.fout[5 * cap + i] = (double)tz;
  }

  // nve/sphere: final integrate of this step (fix_nve_sphere.cpp:208-246) and
  // initial integrate of the next (fix_nve_sphere.cpp:138-206)
  if (a.do_final || a.do_initial) {
    double x = pi.x, y = pi.y, z = pi.z;
    double vx = vi.x, vy = vi.y, vz = vi.z;
    const double dtfm = p.dtf / (vi.w * 1.0);
    const double dtirotate = p.dtfrotate / (pi.w * pi.w * vi.w);
    const double Fx = (double)fx, Fy = (double)fy, Fz = (double)fz;
    T wx = (T)wi.x, wy = (T)wi.y, wz = (T)wi.z;
    const T dtir = (T)dtirotate;
    if (a.do_final) {
      vx += dtfm * Fx;
      vy += dtfm * Fy;
      vz += dtfm * Fz;
      wx += dtir * tx;
      wy += dtir * ty;
      wz += dtir * tz;
    }
    if (a.do_initial) {
      vx += dtfm * Fx;
      vy += dtfm * Fy;
      vz += dtfm * Fz;
      x += p.dt * vx;
      y += p.dt * vy;
      z += p.dt * vz;
      wx += dtir * tx;
      wy += dtir * ty;
      wz += dtir * tz;
      // Neighbor::decide (neighbor.cpp:1366-1371) with delay 0, and
      // Neighbor::check_distance (neighbor.cpp:1436-1443)
      if (ago_next % p.neigh_every == 0) {
        if (p.neigh_check) {
          const double hx = x - a.xhold[i];
          const double hy = y - a.xhold[cap + i];
          const double hz = z - a.xhold[2 * cap + i];
          const double rsq = hx * hx + hy * hy + hz * hz;
          // the lane may hold several tiles (persistent grid): keep a
          // request an earlier tile made
          displaced = displaced || rsq > p.triggersq;
        } else {
          displaced = true;
        }
      }
    }
    a.out.pos[i] = make_double4(x, y, z, pi.w);
    a.out.vel[i] = make_double4(vx, vy, vz, vi.w);
    omg_t wo;
    wo.x = wx;
    wo.y = wy;
    wo.z = wz;
    wo.w = wi.w;
    ((omg_t *)a.out.omg)[i] = wo;
  }
  if (meta != meta0) a.aux.meta[i] = meta;
}

// ---------------------------------------------------------------------------
// fused step kernel: pair forces over the full list, plane walls, meshes,
// gravity, nve/sphere final integrate of this step and initial integrate of
// the next, and the skin/2 displacement check of Neighbor::decide for the
// next step. One thread per particle in tiles of BLOCK particles; the
// first pblocks blocks walk the tiles (one tile each, or several when the
// grid is made persistent). With moving meshes, trailing blocks rotate the
// meshes into the next step's geometry buffer. Every tile shape computes
// the same arithmetic in the same order, so results do not depend on it.
// ---------------------------------------------------------------------------
template <typename T, bool MESH, int BLOCK, int GROUP, bool GEN = false, int WAVE = 0>
__global__ __launch_bounds__(BLOCK) void k_step(StepArgs a) {
  // in a captured graph the step index is the control block's base plus this
  // launch's offset (a.step)
  long long step = a.step, ago_next = a.ago_next;
  if (a.step_from_ctrl) {
    step = a.ctrl->step_base + a.step;
    ago_next = a.ctrl->ago_base + a.step;
  }
  if (step >= a.ctrl->stop_at) return;
  if (MESH && (int)blockIdx.x >= a.pblocks) {
    mesh_move_block(a.mv, ((int)blockIdx.x - a.pblocks) * BLOCK + threadIdx.x, step + 1);
    return;
  }
  const DevParams &p = a.prm;
  typedef typename PairLdsT<T, GROUP, WAVE>::type Lds;
  constexpr bool kWaveLds = WAVE == 1 || WAVE == 2;
  __shared__ Lds lds_all[kWaveLds ? BLOCK / 64 : 1];
  Lds *lds = &lds_all[kWaveLds ? (threadIdx.x >> 6) : 0];
  unsigned long long my_contacts = 0, my_wall = 0;
  bool displaced = false;
  const int ntiles = (a.n + BLOCK - 1) / BLOCK;
  for (int tile = xcd_block(a.pblocks); tile < ntiles; tile += a.pblocks) {
    const int i = tile * BLOCK + threadIdx.x;
    const bool active = i < a.n;
    if (kWaveLds || active) {
      unsigned long long c = 0, w = 0;
      int inst = 0;
      step_particle<T, MESH, GROUP, GEN, WAVE>(a, i, active, c, w, inst, displaced, ago_next, lds);
      my_contacts += c;
      my_wall += w;
      if (active && a.count && a.inst_stats) {
        if (c) atomicAdd(&a.inst_stats[inst], c);
        if (w) atomicAdd(&a.inst_stats[p.ninst + inst], w);
      }
    }
  }

  // rebuild request for the next step
  if (__ballot(displaced) != 0ull) {
    if ((threadIdx.x & 63) == 0)
      atomicMin((unsigned long long *)&a.ctrl->stop_at, (unsigned long long)(step + 1));
  }
  if (a.count) {
    __shared__ unsigned long long sc[BLOCK / 64], sw[BLOCK / 64];
    unsigned long long c = my_contacts, w = my_wall;
    for (int off = 32; off > 0; off >>= 1) {
      c += __shfl_xor(c, off, 64);
      w += __shfl_xor(w, off, 64);
    }
    if ((threadIdx.x & 63) == 0) {
      sc[threadIdx.x >> 6] = c;
      sw[threadIdx.x >> 6] = w;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      unsigned long long tc = 0, tw = 0;
      for (int k = 0; k < BLOCK / 64; k++) {
        tc += sc[k];
        tw += sw[k];
      }
      if (tc) atomicAdd(&a.ctrl->ncontact_full, tc);
      if (tw) atomicAdd(&a.ctrl->nwall, tw);
    }
  }
}

// compute pair/gran/local (compute_pair_gran_local.cpp:312-470) through
// PairGran::compute_pgl (pair_gran.cpp:676-708): one pass over the half list
// with shearupdate off that records, per pair, the force and torque the
// contact puts on its first particle and the pair's stored history
// (PairGran::cpl_add_pair, pair_gran.cpp:798-813, is called for every pair of
// the list; pairs that do not touch get zero force). Each half-list pair is
// taken from the row of its lower-tag particle, whose history entry is
// oriented from that particle to its partner. Count pass, then fill pass at
// the scanned offsets: a particle's pairs are consecutive, in list order.
__global__ __launch_bounds__(kBlock) void k_pair_local_count(StepArgs a, PairLocalArgs o) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= a.n) return;
  const int cnt = a.aux.meta[i] & kMetaCountMask;
  const int ti = a.aux.tag[i];
  int c = 0;
  for (int k = 0; k < cnt; k++) {
    const int j = (int)(a.list.nbr[(size_t)k * a.cap + i] & kIndexMask);
    if (ti < a.aux.tag[j]) c++;
  }
  o.count[i] = c;
}

template <typename T, bool GEN>
__global__ __launch_bounds__(kBlock) void k_pair_local_fill(StepArgs a, PairLocalArgs o) {
  typedef typename OmgT<T>::type omg_t;
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= a.n) return;
  const DevParams &p = a.prm;
  const size_t cap = (size_t)a.cap;
  const size_t hw = (size_t)p.hist_w;
  const double4 pi = a.in.pos[i];
  const double4 vi = a.in.vel[i];
  const omg_t wi = ((const omg_t *)a.in.omg)[i];
  const int wbits = (int)wi.w;
  const int irow = (wbits >> kTypeBits) * p.mat_stride + (wbits & kTypeMask) * p.ntp1;
  const Model pm = model_from_bits(p.pair_model);
  const int cnt = a.aux.meta[i] & kMetaCountMask;
  const int ti = a.aux.tag[i];
  const T *hist = (const T *)a.list.hist;
  int64_t r = o.offset[i];
  for (int k = 0; k < cnt; k++) {
    const uint32_t e = a.list.nbr[(size_t)k * cap + i];
    const int j = (int)(e & kIndexMask);
    const int tj = a.aux.tag[j];
    if (!(ti < tj)) continue;
    const double4 pj = a.in.pos[j];
    const double4 vj = a.in.vel[j];
    double dx, dy, dz;
    min_image(p, pi.x, pi.y, pi.z, pj.x, pj.y, pj.z, dx, dy, dz);
    const double rsq = dx * dx + dy * dy + dz * dz;
    const double radsum = pi.w + pj.w;
    T dF[3] = {T(0), T(0), T(0)}, dTi[3] = {T(0), T(0), T(0)};
    if (rsq < radsum * radsum) pair_contact<T, GEN, omg_t, false>(a, pm, i, k, e, pi, vi, wi, irow, dF, dTi);
    double *rec = o.rec + (size_t)r * (size_t)o.ncol;
    rec[0] = pi.x;
    rec[1] = pi.y;
    rec[2] = pi.z;
    rec[3] = pj.x;
    rec[4] = pj.y;
    rec[5] = pj.z;
    rec[6] = vi.x;
    rec[7] = vi.y;
    rec[8] = vi.z;
    rec[9] = vj.x;
    rec[10] = vj.y;
    rec[11] = vj.z;
    for (int d = 0; d < 3; d++) {
      rec[12 + d] = (double)dF[d];
      rec[15 + d] = (double)dTi[d];
    }
    // the pair crosses a periodic boundary (its partner is a periodic image)
    rec[18] = (dx != pi.x - pj.x || dy != pi.y - pj.y || dz != pi.z - pj.z) ? 1.0 : 0.0;
    rec[19] = pi.w;
    rec[20] = pj.w;
    const bool flagged = (e & kFlagBit) != 0;
    for (size_t d = 0; d < hw; d++)
      rec[kPairLocalFixedCols + d] =
          flagged ? (double)hist[hist_at(k, i, cap, hist_rec((int)hw)) + hist_off((int)d)] : 0.0;
    o.tags[2 * r] = ti;
    o.tags[2 * r + 1] = tj;
    r++;
  }
}

// compute wall/gran/local rows of particle i: the contacts that
// FixWallGran::post_force_pgl (fix_wall_gran.cpp:752-765, computeflag 0,
// shearupdate 0) hands to ComputePairGranLocal::add_wall_1 / add_wall_2
// (fix_wall_gran_base.h:209-219, 325-337): every candidate triangle within
// contact or the close band that handleContact accepts, mesh by mesh. The
// slot table is evaluated on a private copy, so nothing is written; at the
// end of a step it reproduces that step's own contact pass. FILL = false
// only counts the rows.
template <typename T, bool GEN, bool FILL>
__device__ int wall_local_rows(const StepArgs &a, const int i, const WallLocalArgs &o, int64_t r) {
  typedef typename OmgT<T>::type omg_t;
  const DevParams &p = a.prm;
  const MeshStep &M = a.mesh;
  const int nc = M.ncand[i];
  if (nc == 0) return 0;
  const size_t cap = (size_t)a.cap;
  const double4 pi = a.in.pos[i];
  const double4 vi = a.in.vel[i];
  const omg_t wi = ((const omg_t *)a.in.omg)[i];
  const int wbits = (int)wi.w;
  const int irow = (wbits >> kTypeBits) * p.mat_stride + (wbits & kTypeMask) * p.ntp1;
  const double px[3] = {pi.x, pi.y, pi.z};
  int rows = 0;
  int c = 0;
  for (int m = 0; m < p.nmesh; m++) {
    int slot[kMeshSlots];
    const int *gslot = a.aux.mslot + (size_t)m * kMeshSlots * cap + i;
    const size_t mhw = GEN ? (size_t)p.hist_w : 3;
    const T *hist = (const T *)a.aux.mhist + (size_t)m * kMeshSlots * mhw * cap + i;
    for (int k = 0; k < kMeshSlots; k++) slot[k] = gslot[(size_t)k * cap];
    const int tend = p.mesh_tri0[m] + p.mesh_ntri[m];
    const Model mdl = model_from_bits(p.mesh_model[m]);
    const int mt = irow + p.mesh_wtype[m];
    unsigned keep = 0u;
    // slots whose history the step's pass would rewrite, with their values
    unsigned over = 0u;
    T lh[kMeshSlots][GEN ? 6 : 3];
    for (; c < nc; c++) {
      const int t = M.cand[(size_t)c * cap + i];
      if (t >= tend) break;
      const TriGeom &tg = M.tri[t];
      Vec3 delta;
      Bary bary;
      int bs;
      const double deltan = tri_sphere(tg, p.mesh_precision, pi.w, px, delta, bary, bs);
      if (deltan > p.cutneighmax) continue;
      if (!(deltan <= 0 || deltan < (p.cdf - 1.0) * pi.w)) continue;
      // handleContact on the private slot table
      int s = -1;
      for (int k = 0; k < kMeshSlots; k++)
        if (slot[k] == t) {
          s = k;
          break;
        }
      bool fresh = false;
      if (s >= 0) {
        keep |= 1u << s;
      } else {
        bool skip = false;
        if (bs == 7)
          for (int k = 0; k < kMeshSlots; k++)
            if (slot[k] >= 0 && slot[k] != t && ((keep >> k) & 1u) && tri_coplanar(M, slot[k], t)) skip = true;
        if (skip) continue;
        for (int k = 0; k < kMeshSlots; k++)
          if (slot[k] == -1) {
            s = k;
            break;
          }
        if (s < 0) continue;  // the step's own pass reported the overflow
        slot[s] = t;
        keep |= 1u << s;
        fresh = true;
      }
      rows++;
      if (!FILL) continue;
      // the shear, then the epsd2 rolling torque
      const int nh = (GEN && mdl.epsd2) ? 6 : 3;
      T sh[6] = {T(0), T(0), T(0), T(0), T(0), T(0)};
      if (fresh) {
        if (bs == 7)  // checkCoplanarContactHistory: the last coplanar slot wins
          for (int k = 0; k < kMeshSlots; k++)
            if (slot[k] >= 0 && slot[k] != t && tri_coplanar(M, slot[k], t))
              for (int d = 0; d < nh; d++)
                sh[d] = ((over >> k) & 1u) ? lh[k][d] : hist[(size_t)(k * mhw + d) * cap];
      } else {
        for (int d = 0; d < nh; d++) sh[d] = ((over >> s) & 1u) ? lh[s][d] : hist[(size_t)(s * mhw + d) * cap];
      }
      const bool band = !(deltan <= 0);
      if (band && !mdl.nohist) sh[0] = sh[1] = sh[2] = T(0);
      if (band) sh[3] = sh[4] = sh[5] = T(0);
      // the step's pass rewrites a fresh or close-band slot with these values
      if (fresh || band) {
        for (int d = 0; d < nh; d++) lh[s][d] = sh[d];
        over |= 1u << s;
      }
      double vw[3] = {0., 0., 0.};
      if (p.mesh_moving[m])
        for (int k = 0; k < 3; k++)
          vw[k] = (bary.b0 * tg.vn[k] + bary.b1 * tg.vn[3 + k] + bary.b2 * tg.vn[6 + k]);
      // Walls::Granular::compute_force (fix_wall_gran_base.h:148-337) with
      // sidata.deltan = -deltan, sidata.delta = -delta
      const double sdn = -deltan;
      const double rr = pi.w - sdn;
      const double rinv = 1.0 / rr;
      const double sd0 = -delta.x, sd1 = -delta.y, sd2 = -delta.z;
      T dF[3] = {T(0), T(0), T(0)}, dTi[3] = {T(0), T(0), T(0)};
      if (!band) {
        CIn<T> cc;
        cc.radi = (T)pi.w;
        cc.radj = T(0);
        cc.rinv = (T)rinv;
        cc.meff = (T)vi.w;
        cc.deltan = (T)(pi.w - rr);
        cc.deltan_in = (T)sdn;
        cc.del[0] = (T)sd0;
        cc.del[1] = (T)sd1;
        cc.del[2] = (T)sd2;
        cc.en[0] = (T)(sd0 * rinv);
        cc.en[1] = (T)(sd1 * rinv);
        cc.en[2] = (T)(sd2 * rinv);
        cc.vr[0] = (T)(vi.x - vw[0]);
        cc.vr[1] = (T)(vi.y - vw[1]);
        cc.vr[2] = (T)(vi.z - vw[2]);
        cc.wi[0] = (T)wi.x;
        cc.wi[1] = (T)wi.y;
        cc.wi[2] = (T)wi.z;
        cc.wj[0] = cc.wj[1] = cc.wj[2] = T(0);
        if (sizeof(T) == 8) {
          cc.Y = (T)a.mat.Yeff[mt];
          cc.G = (T)a.mat.Geff[mt];
          cc.beta = (T)a.mat.beta[mt];
          cc.mu = (T)a.mat.mu[mt];
          cc.mur = (T)a.mat.mur[mt];
        } else {
          cc.Y = (T)a.mat.fYeff[mt];
          cc.G = (T)a.mat.fGeff[mt];
          cc.beta = (T)a.mat.fbeta[mt];
          cc.mu = (T)a.mat.fmu[mt];
          cc.mur = (T)a.mat.fmur[mt];
        }
        if (GEN) {
          load_model_coeffs(a.mat, p, mdl, mt, cc);
          cc.acont = (T)((pi.w * pi.w - rr * rr) * kPi);
        }
        T dTj[3];
        contact_chain<T, true, GEN>(mdl, cc, (T)p.dt, 0, sh, sh + 3, dF, dTi, dTj);
      }
      double *rec = o.rec + (size_t)(r + rows - 1) * (size_t)o.ncol;
      // contact point x - sidata.delta (fix_wall_gran_base.h:211-212)
      rec[0] = px[0] + delta.x;
      rec[1] = px[1] + delta.y;
      rec[2] = px[2] + delta.z;
      rec[3] = px[0];
      rec[4] = px[1];
      rec[5] = px[2];
      rec[6] = vw[0];
      rec[7] = vw[1];
      rec[8] = vw[2];
      rec[9] = vi.x;
      rec[10] = vi.y;
      rec[11] = vi.z;
      for (int d = 0; d < 3; d++) {
        rec[12 + d] = (double)dF[d];
        rec[15 + d] = (double)dTi[d];
      }
      // normal = -sidata.en = delta / r (fix_wall_gran_base.h:332-334), in FP64
      rec[18] = -sd0 * rinv;
      rec[19] = -sd1 * rinv;
      rec[20] = -sd2 * rinv;
      rec[21] = rr * rr;
      rec[22] = pi.w;
      for (int d = 0; d < 3; d++) rec[kWallLocalFixedCols + d] = (double)sh[d];
      int *id = o.ids + 3 * (size_t)(r + rows - 1);
      id[0] = m;
      id[1] = t - p.mesh_tri0[m];
      id[2] = a.aux.tag[i];
    }
  }
  return rows;
}

template <typename T, bool GEN>
__global__ __launch_bounds__(kBlock) void k_wall_local_count(StepArgs a, WallLocalArgs o) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= a.n) return;
  o.count[i] = wall_local_rows<T, GEN, false>(a, i, o, 0);
}

template <typename T, bool GEN>
__global__ __launch_bounds__(kBlock) void k_wall_local_fill(StepArgs a, WallLocalArgs o) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= a.n) return;
  wall_local_rows<T, GEN, true>(a, i, o, o.offset[i]);
}

// initial integrate of step 1 after setup (the setup force pass has no
// final integrate), plus the displacement check for step 1. In place.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_initial(StepArgs a) {
  typedef typename OmgT<T>::type omg_t;
  const int i = blockIdx.x * kBlock + threadIdx.x;
  bool displaced = false;
  if (i < a.n) {
    const DevParams &p = a.prm;
    const size_t cap = (size_t)a.cap;
    const double4 pi = a.in.pos[i];
    const double4 vi = a.in.vel[i];
    omg_t wi = ((omg_t *)a.in.omg)[i];
    const double dtfm = p.dtf / (vi.w * 1.0);
    const double dtirotate = p.dtfrotate / (pi.w * pi.w * vi.w);
    const T dtir = (T)dtirotate;
    const double vx = vi.x + dtfm * a.fout[i];
    const double vy = vi.y + dtfm * a.fout[cap + i];
    const double vz = vi.z + dtfm * a.fout[2 * cap + i];
    const double x = pi.x + p.dt * vx, y = pi.y + p.dt * vy, z = pi.z + p.dt * vz;
    wi.x += dtir * (T)a.fout[3 * cap + i];
    wi.y += dtir * (T)a.fout[4 * cap + i];
    wi.z += dtir * (T)a.fout[5 * cap + i];
    a.in.pos[i] = make_double4(x, y, z, pi.w);
    a.in.vel[i] = make_double4(vx, vy, vz, vi.w);
    ((omg_t *)a.in.omg)[i] = wi;
    if (a.ago_next % p.neigh_every == 0) {
      if (p.neigh_check) {
        const double hx = x - a.xhold[i];
        const double hy = y - a.xhold[cap + i];
        const double hz = z - a.xhold[2 * cap + i];
        displaced = hx * hx + hy * hy + hz * hz > p.triggersq;
      } else {
        displaced = true;
      }
    }
  }
  if (__ballot(displaced) != 0ull && (threadIdx.x & 63) == 0)
    atomicMin((unsigned long long *)&a.ctrl->stop_at, 1ull);
}

__global__ void k_set_base(Ctrl *ctrl, long long step, long long ago) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    ctrl->step_base = step;
    ctrl->ago_base = ago;
  }
}

__global__ void k_set_stop(Ctrl *ctrl, long long v, unsigned long long *inst_stats, int nstat) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k == 0) {
    ctrl->stop_at = v;
    ctrl->ncontact_full = 0;
    ctrl->nwall = 0;
  }
  if (k < kMaxMeshes * 6) ctrl->mesh_force[k / 6][k % 6] = 0.0;
  if (inst_stats && k < nstat) inst_stats[k] = 0ull;
}

// ---------------------------------------------------------------------------
// rebuild: Domain::pbc (domain.cpp:530-620) + Morton-cell counting sort +
// gather + full neighbor list with history carry-in (neigh_gran.cpp:485-639,
// fix_contact_history.cpp:305-425)
// ---------------------------------------------------------------------------
// first global cell of the instance particle i belongs to (omega.w holds
// type + instance << kTypeBits)
__device__ __forceinline__ int inst_cell_base(const DevParams &p, const StateBuf &s, int mixed,
                                              int i) {
  if (p.ninst <= 1) return 0;
  const int w = mixed ? (int)((const float4 *)s.omg)[i].w : (int)((const double4 *)s.omg)[i].w;
  return (w >> kTypeBits) * p.ncell_local;
}

__global__ __launch_bounds__(kBlock) void k_wrap_keys(RebuildArgs a) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= a.n) return;
  const DevParams &p = a.prm;
  double4 q = a.cur.pos[i];
  double xs[3] = {q.x, q.y, q.z};
  bool changed = false;
  for (int d = 0; d < 3; d++) {
    if (!p.per[d]) continue;
    const double lo = p.lo[d], hi = p.lo[d] + p.L[d];
    if (xs[d] < lo) {
      xs[d] += p.L[d];
      changed = true;
    }
    if (xs[d] >= hi) {
      xs[d] -= p.L[d];
      if (xs[d] < lo) xs[d] = lo;
      changed = true;
    }
  }
  if (changed) {
    q.x = xs[0];
    q.y = xs[1];
    q.z = xs[2];
    a.cur.pos[i] = q;
  }
  // Comm::exchange (comm.cpp:945-1003) with one process: an atom with
  // x < lo or x >= hi in a dimension is sent to itself and taken back only
  // inside the box, so beyond an f face it is lost
  bool lost = false;
  for (int d = 0; d < 3; d++) {
    if (p.per[d]) continue;
    const double lo = p.lo[d], hi = p.lo[d] + p.L[d];
    if (((p.lost_faces >> (2 * d)) & 1) && xs[d] < lo) lost = true;
    if (((p.lost_faces >> (2 * d + 1)) & 1) && xs[d] >= hi) lost = true;
  }
  if (lost) atomicAdd(&a.ctrl->nlost, 1ull);
  const int cx = cell_coord(p, 0, xs[0]);
  const int cy = cell_coord(p, 1, xs[1]);
  const int cz = cell_coord(p, 2, xs[2]);
  const int lin = cx + p.nc[0] * (cy + p.nc[1] * cz);
  const int key = inst_cell_base(p, a.cur, a.mixed, i) + a.cell_rank[lin];
  a.key[i] = key;
  atomicAdd(&a.cell_count[key], 1);
}

// exclusive scan, three passes (block scan, scan of block sums, add)
constexpr int kScanPer = 1024;  // elements per scan block (256 threads x 4)

__global__ __launch_bounds__(256) void k_scan_block(const int *in, int *out, int *bsum, int n) {
  __shared__ int s[256];
  const int base = blockIdx.x * kScanPer + threadIdx.x * 4;
  int v[4];
  int t = 0;
  for (int k = 0; k < 4; k++) {
    v[k] = (base + k < n) ? in[base + k] : 0;
    t += v[k];
  }
  s[threadIdx.x] = t;
  __syncthreads();
  for (int off = 1; off < 256; off <<= 1) {
    const int add = threadIdx.x >= off ? s[threadIdx.x - off] : 0;
    __syncthreads();
    s[threadIdx.x] += add;
    __syncthreads();
  }
  int run = s[threadIdx.x] - t;  // exclusive prefix of this thread
  for (int k = 0; k < 4; k++) {
    if (base + k < n) out[base + k] = run;
    run += v[k];
  }
  if (threadIdx.x == 255) bsum[blockIdx.x] = s[255];
}

__global__ __launch_bounds__(256) void k_scan_top(int *bsum, int nb, int *total) {
  __shared__ int s[256];
  __shared__ int carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (int base = 0; base < nb; base += 256) {
    const int idx = base + threadIdx.x;
    const int v = idx < nb ? bsum[idx] : 0;
    s[threadIdx.x] = v;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
      const int add = threadIdx.x >= off ? s[threadIdx.x - off] : 0;
      __syncthreads();
      s[threadIdx.x] += add;
      __syncthreads();
    }
    if (idx < nb) bsum[idx] = carry + s[threadIdx.x] - v;
    __syncthreads();
    if (threadIdx.x == 255) carry += s[255];
    __syncthreads();
  }
  if (threadIdx.x == 0) *total = carry;
}

__global__ __launch_bounds__(256) void k_scan_add(int *out, const int *bsum, int n) {
  const int base = blockIdx.x * kScanPer + threadIdx.x * 4;
  const int add = bsum[blockIdx.x];
  for (int k = 0; k < 4; k++)
    if (base + k < n) out[base + k] += add;
}

__global__ __launch_bounds__(kBlock) void k_zero(int *p, int n) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) p[i] = 0;
}

__global__ __launch_bounds__(kBlock) void k_scatter(RebuildArgs a) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= a.n) return;
  const int key = a.key[i];
  const int slot = a.cell_start[key] + atomicAdd(&a.cell_count[key], 1);
  a.perm[slot] = i;
}

// order the particles of each cell by tag, so the layout is deterministic
__global__ __launch_bounds__(kBlock) void k_cell_sort(RebuildArgs a) {
  const int c = blockIdx.x * kBlock + threadIdx.x;
  if (c >= a.ncell) return;
  const int s = a.cell_start[c], e = a.cell_start[c + 1];
  for (int k = s + 1; k < e; k++) {
    const int v = a.perm[k];
    const int tv = a.acur.tag[v];
    int m = k - 1;
    while (m >= s && a.acur.tag[a.perm[m]] > tv) {
      a.perm[m + 1] = a.perm[m];
      m--;
    }
    a.perm[m + 1] = v;
  }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_gather(RebuildArgs a) {
  typedef typename OmgT<T>::type omg_t;
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= a.n) return;
  const int o = a.perm[i];
  const size_t cap = (size_t)a.cap;
  const double4 q = a.cur.pos[o];
  a.alt.pos[i] = q;
  a.alt.vel[i] = a.cur.vel[o];
  ((omg_t *)a.alt.omg)[i] = ((const omg_t *)a.cur.omg)[o];
  a.aalt.tag[i] = a.acur.tag[o];
  a.aalt.orig[i] = a.acur.orig[o];
  const int meta = a.acur.meta[o];
  a.aalt.meta[i] = meta & ~kMetaCountMask;
  const T *wi = (const T *)a.acur.whist;
  T *wo = (T *)a.aalt.whist;
  const int hw = a.prm.hist_w;
  for (int w = 0; w < a.prm.nwalls; w++) {
    if (!(meta & (1 << (kMetaWallShift + w)))) continue;
    for (int d = 0; d < hw; d++) wo[(size_t)(w * hw + d) * cap + i] = wi[(size_t)(w * hw + d) * cap + o];
  }
  a.xhold[i] = q.x;
  a.xhold[cap + i] = q.y;
  a.xhold[2 * cap + i] = q.z;
  const T *hi = (const T *)a.acur.mhist;
  T *ho = (T *)a.aalt.mhist;
  for (int m = 0; m < a.prm.nmesh; m++)
    for (int k = 0; k < kMeshSlots; k++) {
      const size_t sb = (size_t)(m * kMeshSlots + k) * cap;
      const int st = a.acur.mslot[sb + o];
      a.aalt.mslot[sb + i] = st;
      if (st >= 0)
        for (int d = 0; d < a.prm.hist_w; d++) {
          const size_t hb = (size_t)((m * kMeshSlots + k) * a.prm.hist_w + d) * cap;
          ho[hb + i] = hi[hb + o];
        }
    }
}

// Flagged partners of a particle's old list row that the build looks up by
// binary search: up to kCarrySlots per particle, in LDS, keyed by the
// partner's index in the old particle order (ascending), with the old list
// slot beside it. A row with more flagged entries falls back to scanning the
// old row in global memory.
constexpr int kCarrySlots = 16;
// List entries per particle that the build stages in LDS before writing
// them row by row; entries past this many go straight to their slot.
constexpr int kStageSlots = 16;

struct BuildLds {
  uint32_t key[kCarrySlots][kBlock];
  uint16_t slot[kCarrySlots][kBlock];
  // staged entries in stencil order
  uint32_t entry[kStageSlots][kBlock];
  // per touching staged entry, by touching rank: its old slot + 1, or 0
  // when it carries no history
  uint16_t carried[kCarrySlots][kBlock];
};

// Loads the flagged entries of old row io (oldcnt entries) into this lane's
// sorted LDS column; returns their number, or -1 when they do not fit.
__device__ __forceinline__ int carry_table(const RebuildArgs &a, const size_t cap, const int io,
                                           const int oldcnt, BuildLds &L, const int t) {
  int nf = 0;
  for (int m0 = 0; m0 < oldcnt; m0 += 4) {
    uint32_t oe[4];
#pragma unroll
    for (int u = 0; u < 4; u++) oe[u] = (m0 + u < oldcnt) ? a.oldl.nbr[(size_t)(m0 + u) * cap + io] : 0u;
#pragma unroll
    for (int u = 0; u < 4; u++) {
      if (!(oe[u] & kFlagBit)) continue;
      if (nf == kCarrySlots) return -1;
      // insertion into the ascending key column
      const uint32_t key = oe[u] & kIndexMask;
      int q = nf;
      while (q > 0 && L.key[q - 1][t] > key) {
        L.key[q][t] = L.key[q - 1][t];
        L.slot[q][t] = L.slot[q - 1][t];
        q--;
      }
      L.key[q][t] = key;
      L.slot[q][t] = (uint16_t)(m0 + u);
      nf++;
    }
  }
  return nf;
}

// Old list slot of flagged partner jo, or -1: binary search of the LDS column
// (nf >= 0) or a scan of the old row (nf < 0).
__device__ __forceinline__ int carry_find(const RebuildArgs &a, const size_t cap, const int io, const int oldcnt,
                                          const BuildLds &L, const int t, const int nf, const uint32_t jo) {
  if (nf >= 0) {
    int lo = 0, hi = nf;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (L.key[mid][t] < jo)
        lo = mid + 1;
      else
        hi = mid;
    }
    return (lo < nf && L.key[lo][t] == jo) ? (int)L.slot[lo][t] : -1;
  }
  for (int m = 0; m < oldcnt; m++) {
    const uint32_t oe = a.oldl.nbr[(size_t)m * cap + io];
    if ((oe & kFlagBit) && (oe & kIndexMask) == jo) return m;
  }
  return -1;
}

template <typename T>
__device__ __forceinline__ void copy_history(const RebuildArgs &a, const size_t cap, const size_t hw,
                                             const int io, const int m, const int i, const int k) {
  const T *ohist = (const T *)a.oldl.hist;
  T *nhist = (T *)a.newl.hist;
  const size_t rec = hist_rec((int)hw);
  const size_t ob = hist_at(m, io, cap, rec);
  const size_t nb = hist_at(k, i, cap, rec);
  T v[3];
  load_rec3(ohist + ob, v);
  store_rec3(nhist + nb, v);
  if (hw == 6) {
    load_rec3(ohist + ob + 4, v);
    store_rec3(nhist + nb + 4, v);
  }
}

// Neighbor::granular_bin_no_newton (neigh_gran.cpp:485-639) as a full list,
// one thread per particle: the reference's predicate rsq <= ((ri+rj)*cdf +
// skin)^2 over the 27-cell stencil; history carry-in for pairs within
// (ri+rj)*cdf (neigh_gran.cpp:585-607: a flagged partner of the old row
// passes its history and flag 1). The candidates of each stencil row (three
// cells in x) are tested in batches of four independent position loads, and
// old partners are found by binary search in LDS.
//
// Row order: a particle's entries within (ri+rj)*cdf (the touching ones)
// come first, then the others, each in stencil order ((z, y, x) cells, j
// ascending within a cell); entries past kStageSlots keep their stencil
// slot. The reference's pair order only sets the order of the force sums.
// Touching-first rows put the step's history reads and writes of a
// wavefront's particles in the first few list rows, side by side.
//
// The first kStageSlots entries and the old slots of the touching ones wait
// in LDS and are written row by row at the end, so the lanes of a wavefront
// write the same list row (and history row) together. Workgroups walk the
// particles XCD by XCD (xcd_block), so a neighborhood is fetched into one L2.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_build(RebuildArgs a) {
  __shared__ BuildLds lds;
  const int t = threadIdx.x;
  const int i = xcd_block(gridDim.x) * kBlock + t;
  const bool active = i < a.n;
  int cnt = 0;
  if (active) {
    const DevParams &p = a.prm;
    const size_t cap = (size_t)a.cap;
    const double4 pi = a.alt.pos[i];
    const int io = a.perm[i];
    const int oldcnt = a.have_old ? (a.acur.meta[io] & kMetaCountMask) : 0;
    const int ci[3] = {cell_coord(p, 0, pi.x), cell_coord(p, 1, pi.y), cell_coord(p, 2, pi.z)};
    const int cbase = inst_cell_base(p, a.alt, a.mixed, i);
    int lo[3], hi[3];
    for (int d = 0; d < 3; d++) {
      if (p.nc[d] == 1) {
        lo[d] = hi[d] = 0;
      } else if (p.per[d]) {
        lo[d] = -1;
        hi[d] = 1;
      } else {
        lo[d] = ci[d] > 0 ? -1 : 0;
        hi[d] = ci[d] < p.nc[d] - 1 ? 1 : 0;
      }
    }
    // cell ranges of the stencil, cell c = (dz+1)*9 + (dy+1)*3 + (dx+1)
    int rk[27];
#pragma unroll
    for (int c = 0; c < 27; c++) {
      const int dz = c / 9 - 1, dy = (c / 3) % 3 - 1, dx = c % 3 - 1;
      const bool ok = dz >= lo[2] && dz <= hi[2] && dy >= lo[1] && dy <= hi[1] && dx >= lo[0] && dx <= hi[0];
      const int cx = (ci[0] + dx + p.nc[0]) % p.nc[0];
      const int cy = (ci[1] + dy + p.nc[1]) % p.nc[1];
      const int cz = (ci[2] + dz + p.nc[2]) % p.nc[2];
      rk[c] = ok ? cbase + a.cell_rank[cx + p.nc[0] * (cy + p.nc[1] * cz)] : -1;
    }
    int cs[27], cn[27];
#pragma unroll
    for (int c = 0; c < 27; c++) {
      const int r = rk[c] < 0 ? 0 : rk[c];
      const int s0 = a.cell_start[r], s1 = a.cell_start[r + 1];
      cs[c] = s0;
      cn[c] = rk[c] < 0 ? 0 : s1 - s0;
    }
    const int nf = oldcnt > 0 ? carry_table(a, cap, io, oldcnt, lds, t) : 0;
    const size_t hw = (size_t)p.hist_w;
    const int stage = a.W < kStageSlots ? a.W : kStageSlots;
    uint32_t tmask = 0u;  // staged entries that touch
    int nts = 0;          // their number
#pragma unroll
    for (int row = 0; row < 9; row++) {
      const int s0 = cs[3 * row], s1 = cs[3 * row + 1], s2 = cs[3 * row + 2];
      const int n0 = cn[3 * row], n01 = n0 + cn[3 * row + 1], nt = n01 + cn[3 * row + 2];
      for (int k0 = 0; k0 < nt; k0 += 4) {
        int jj[4];
        double4 pj[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
          const int k = k0 + u;
          const int j = k < n0 ? s0 + k : (k < n01 ? s1 + (k - n0) : s2 + (k - n01));
          jj[u] = k < nt ? j : i;
        }
#pragma unroll
        for (int u = 0; u < 4; u++) pj[u] = a.alt.pos[jj[u]];
        double ddx[4], ddy[4], ddz[4];
        bool wrap = false;
#pragma unroll
        for (int u = 0; u < 4; u++) {
          ddx[u] = pi.x - pj[u].x;
          ddy[u] = pi.y - pj[u].y;
          ddz[u] = pi.z - pj[u].z;
          wrap |= needs_image(p, ddx[u], ddy[u], ddz[u]);
        }
        if (wrap) {
#pragma unroll
          for (int u = 0; u < 4; u++) apply_image(p, ddx[u], ddy[u], ddz[u]);
        }
#pragma unroll
        for (int u = 0; u < 4; u++) {
          const int j = jj[u];
          if (j == i) continue;  // i itself, or past the row's candidates
          const double rsq = ddx[u] * ddx[u] + ddy[u] * ddy[u] + ddz[u] * ddz[u];
          const double radsum = (pi.w + pj[u].w) * p.cdf;
          const double cutsq = (radsum + p.skin) * (radsum + p.skin);
          if (!(rsq <= cutsq)) continue;
          if (cnt < a.W) {
            uint32_t entry = (uint32_t)j;
            const bool touch = rsq < radsum * radsum;
            int m = -1;
            if (touch && oldcnt > 0) {
              m = carry_find(a, cap, io, oldcnt, lds, t, nf, (uint32_t)a.perm[j]);
              if (m >= 0) entry |= kFlagBit | kCarryBit;
            }
            if (cnt < stage) {
              lds.entry[cnt][t] = entry;
              if (touch) {
                if (nts < kCarrySlots)
                  lds.carried[nts][t] = (uint16_t)(m + 1);
                else if (m >= 0)
                  copy_history<T>(a, cap, hw, io, m, i, nts);
                tmask |= 1u << cnt;
                nts++;
              }
            } else {
              a.newl.nbr[(size_t)cnt * cap + i] = entry;
              if (m >= 0) copy_history<T>(a, cap, hw, io, m, i, cnt);
            }
          }
          cnt++;
        }
      }
    }
    // staged entries, touching ones first, one list row at a time
    const int staged = cnt < stage ? cnt : stage;
    uint32_t mt = tmask, mn = ~tmask & (staged < 32 ? (1u << staged) - 1u : ~0u);
    for (int s = 0; s < staged; s++) {
      int k;
      if (mt) {
        k = __builtin_ctz(mt);
        mt &= mt - 1u;
      } else {
        k = __builtin_ctz(mn);
        mn &= mn - 1u;
      }
      a.newl.nbr[(size_t)s * cap + i] = lds.entry[k][t];
    }
    // histories of the carried staged entries, at their touching rank
    const int nq = nts < kCarrySlots ? nts : kCarrySlots;
    for (int q = 0; q < nq; q++) {
      const int m1 = lds.carried[q][t];
      if (m1) copy_history<T>(a, cap, hw, io, m1 - 1, i, q);
    }
    const int stored = cnt < a.W ? cnt : a.W;
    a.aalt.meta[i] = (a.aalt.meta[i] & ~kMetaCountMask) | stored;
    if (cnt > a.W) atomicMax(&a.ctrl->overflow, cnt);
  }
  // widest list and entry count, one atomic per wavefront
  int wmax = cnt;
  unsigned long long wsum = (unsigned long long)cnt;
  for (int off = 32; off > 0; off >>= 1) {
    const int o = __shfl_xor(wmax, off, 64);
    wmax = o > wmax ? o : wmax;
    wsum += __shfl_xor(wsum, off, 64);
  }
  if ((threadIdx.x & 63) == 0 && wsum) {
    atomicMax(&a.ctrl->maxcount, wmax);
    atomicAdd(&a.ctrl->npairs_full, wsum);
  }
}

__global__ void k_reset_build_counters(Ctrl *ctrl) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    ctrl->overflow = 0;
    ctrl->maxcount = 0;
    ctrl->npairs_full = 0;
    ctrl->mesh_overflow = 0;
    ctrl->nmesh_cand = 0;
    ctrl->nlost = 0;
  }
}

// ---------------------------------------------------------------------------
// triangle meshes: rebuild-time geometry refresh, binning and candidates
// ---------------------------------------------------------------------------

// FixMesh::pre_force on a rebuild step refreshes the element geometry from the
// nodes and stores the node positions (pbcExchangeBorders -> refreshOwned,
// multi_node_mesh_I.h:212-226, surface_mesh_I.h:146-200)
__global__ __launch_bounds__(kBlock) void k_mesh_geom(RebuildArgs a) {
  const int t = blockIdx.x * kBlock + threadIdx.x;
  if (t >= a.ntri) return;
  TriGeom &tg = a.tri[t];
  tri_surf_props(tg);
  for (int k = 0; k < 9; k++) a.node_last[9 * t + k] = tg.n[k];
  if (a.zero_vnode)
    for (int k = 0; k < 9; k++) tg.vn[k] = 0.0;
  // the other buffer holds the same geometry until a step moves it
  if (a.tri_alt) a.tri_alt[t] = tg;
}

// cells a triangle can reach: centre within (r_max*cdf + skin) + half a cell
// diagonal of the triangle
__device__ __forceinline__ void tri_cell_range(const RebuildArgs &a, const TriGeom &tg, int lo[3],
                                               int hi[3], double &reach2) {
  const DevParams &p = a.prm;
  const double e = p.cutneighmax - p.skin;  // (r_max + r_max) * cdf
  const double reach = 0.5 * e + p.skin;    // r_max * cdf + skin
  double hd = 0.0;
  for (int d = 0; d < 3; d++) {
    double mn = tg.n[d], mx = tg.n[d];
    for (int j = 1; j < 3; j++) {
      mn = fmin(mn, tg.n[3 * j + d]);
      mx = fmax(mx, tg.n[3 * j + d]);
    }
    lo[d] = cell_coord(p, d, mn - reach);
    hi[d] = cell_coord(p, d, mx + reach);
    const double cs = p.L[d] / p.nc[d];
    hd += 0.25 * cs * cs;
  }
  const double r = reach + sqrt(hd);
  reach2 = r * r;
}

__device__ __forceinline__ void cell_center(const DevParams &p, int cx, int cy, int cz, double *c) {
  c[0] = p.lo[0] + (cx + 0.5) * (p.L[0] / p.nc[0]);
  c[1] = p.lo[1] + (cy + 0.5) * (p.L[1] / p.nc[1]);
  c[2] = p.lo[2] + (cz + 0.5) * (p.L[2] / p.nc[2]);
}

// Triangle -> cell binning, one wavefront per triangle: the lanes stride over
// the cells of the triangle's grown bounding box and keep those whose centre
// lies within reach (+ half the cell diagonal) of the triangle.
constexpr int kWavesPerBlock = kBlock / 64;

__device__ __forceinline__ bool tri_cell_hit(const RebuildArgs &a, const TriGeom &tg,
                                             const int lo[3], const int nx, const int ny,
                                             const double reach2, const int k, int &cell) {
  const int cx = lo[0] + k % nx;
  const int cy = lo[1] + (k / nx) % ny;
  const int cz = lo[2] + k / (nx * ny);
  double cc[3];
  cell_center(a.prm, cx, cy, cz, cc);
  cell = cx + a.prm.nc[0] * (cy + a.prm.nc[1] * cz);
  return tri_point_dist2(&tg.n[0], &tg.n[3], &tg.n[6], cc) <= reach2;
}

__global__ __launch_bounds__(kBlock) void k_tri_count(RebuildArgs a) {
  const int t = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (t >= a.ntri) return;  // uniform per wavefront
  const TriGeom &tg = a.tri[t];
  int lo[3], hi[3];
  double reach2;
  tri_cell_range(a, tg, lo, hi, reach2);
  const int nx = hi[0] - lo[0] + 1, ny = hi[1] - lo[1] + 1, nz = hi[2] - lo[2] + 1;
  const int ncells = nx * ny * nz;
  int cnt = 0;
  for (int k = lane; k < ncells; k += 64) {
    int cell;
    cnt += tri_cell_hit(a, tg, lo, nx, ny, reach2, k, cell) ? 1 : 0;
  }
  for (int off = 32; off > 0; off >>= 1) cnt += __shfl_xor(cnt, off, 64);
  if (lane == 0) a.tcount[t] = cnt;
}

__global__ __launch_bounds__(kBlock) void k_tri_fill(RebuildArgs a) {
  const int t = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (t >= a.ntri) return;  // uniform per wavefront
  const TriGeom &tg = a.tri[t];
  int lo[3], hi[3];
  double reach2;
  tri_cell_range(a, tg, lo, hi, reach2);
  const int nx = hi[0] - lo[0] + 1, ny = hi[1] - lo[1] + 1, nz = hi[2] - lo[2] + 1;
  const int ncells = nx * ny * nz;
  int w = a.toff[t];
  const int wend = w + a.tcount[t];
  const unsigned long long below = (1ull << lane) - 1ull;
  for (int k0 = 0; k0 < ncells; k0 += 64) {  // uniform trip count
    const int k = k0 + lane;
    int cell = 0;
    const bool hit = k < ncells && tri_cell_hit(a, tg, lo, nx, ny, reach2, k, cell);
    const unsigned long long m = __ballot(hit);
    if (hit) {
      const int pos = w + __popcll(m & below);
      if (pos < wend) {
        a.tpair_cell[pos] = a.cell_rank[cell];
        a.tpair_tri[pos] = t;
      }
    }
    w += __popcll(m);
  }
}

// (triangle, cell) pairs hold instance-local cell ranks; every instance sees
// the same walls, so pair k of instance q lands in global cell
// q * ncell_local + rank. k runs over npairs * ninst.
__global__ __launch_bounds__(kBlock) void k_tcell_hist(RebuildArgs a, int npairs) {
  const int k = blockIdx.x * kBlock + threadIdx.x;
  if (k >= npairs * a.prm.ninst) return;
  const int q = k / npairs, pr = k - q * npairs;
  atomicAdd(&a.tcell_count[q * a.prm.ncell_local + a.tpair_cell[pr]], 1);
}

__global__ __launch_bounds__(kBlock) void k_tcell_scatter(RebuildArgs a, int npairs) {
  const int k = blockIdx.x * kBlock + threadIdx.x;
  if (k >= npairs * a.prm.ninst) return;
  const int q = k / npairs, pr = k - q * npairs;
  const int c = q * a.prm.ncell_local + a.tpair_cell[pr];
  const int pos = a.tcell_start[c] + atomicAdd(&a.tcell_count[c], 1);
  a.tcell_list[pos] = a.tpair_tri[pr];
}

template <typename T>
__device__ __forceinline__ void slot_swap(int *slot, T *hist, size_t cap, int hw, int x, int y) {
  const int ts = slot[(size_t)x * cap];
  slot[(size_t)x * cap] = slot[(size_t)y * cap];
  slot[(size_t)y * cap] = ts;
  for (int d = 0; d < hw; d++) {
    const T h = hist[(size_t)(x * hw + d) * cap];
    hist[(size_t)(x * hw + d) * cap] = hist[(size_t)(y * hw + d) * cap];
    hist[(size_t)(y * hw + d) * cap] = h;
  }
}

// per particle: FixContactHistoryMesh::sort_contacts (slot compaction,
// fix_contact_history_mesh.cpp:374-400), candidate triangles within
// r*cdf + skin in ascending index, and cleanUpContactJumps (:466-500)
template <typename T>
__device__ int mesh_cand_one(const RebuildArgs &a, const int i) {
  const DevParams &p = a.prm;
  const size_t cap = (size_t)a.cap;
  const double4 pi = a.alt.pos[i];
  // candidates
  const int cx = cell_coord(p, 0, pi.x), cy = cell_coord(p, 1, pi.y), cz = cell_coord(p, 2, pi.z);
  const int rank = inst_cell_base(p, a.alt, a.mixed, i) + a.cell_rank[cx + p.nc[0] * (cy + p.nc[1] * cz)];
  const double lim = pi.w * p.cdf + p.skin;
  const double px[3] = {pi.x, pi.y, pi.z};
  int nc = 0;
  for (int k = a.tcell_start[rank]; k < a.tcell_start[rank + 1]; k++) {
    const int t = a.tcell_list[k];
    const TriGeom &tg = a.tri[t];
    if (tri_point_dist2(&tg.n[0], &tg.n[3], &tg.n[6], px) <= lim * lim) {
      if (nc < kMaxMeshCand) {
        // insertion into the ascending list
        int m = nc - 1;
        while (m >= 0 && a.cand[(size_t)m * cap + i] > t) {
          a.cand[(size_t)(m + 1) * cap + i] = a.cand[(size_t)m * cap + i];
          m--;
        }
        a.cand[(size_t)(m + 1) * cap + i] = t;
      }
      nc++;
    }
  }
  if (nc > kMaxMeshCand) {
    atomicMax(&a.ctrl->mesh_overflow, nc);
    nc = kMaxMeshCand;
  }
  a.ncand[i] = nc;
  // slots, mesh by mesh
  for (int m = 0; m < p.nmesh; m++) {
    int *slot = a.aalt.mslot + (size_t)m * kMeshSlots * cap + i;
    const int hw = p.hist_w;
    T *hist = (T *)a.aalt.mhist + (size_t)m * kMeshSlots * hw * cap + i;
    int first_empty, last_filled;
    do {
      first_empty = last_filled = -1;
      for (int j = 0; j < kMeshSlots; j++) {
        const int st = slot[(size_t)j * cap];
        if (-1 == first_empty && -1 == st) first_empty = j;
        if (st >= 0) last_filled = j;
      }
      if (first_empty > -1 && last_filled > -1 && first_empty < last_filled)
        slot_swap(slot, hist, cap, hw, first_empty, last_filled);
    } while (first_empty > -1 && last_filled > -1 && first_empty < last_filled);
    int np = 0;
    for (int j = 0; j < kMeshSlots; j++) np += slot[(size_t)j * cap] >= 0;
    int ip = 0;
    while (ip < np) {
      const int st = slot[(size_t)ip * cap];
      bool inlist = false;
      for (int k = 0; k < nc; k++)
        if (a.cand[(size_t)k * cap + i] == st) inlist = true;
      if (!inlist) {
        slot[(size_t)ip * cap] = -1;
        for (int d = 0; d < hw; d++) hist[(size_t)(ip * hw + d) * cap] = T(0);
        slot_swap(slot, hist, cap, hw, ip, np - 1);
        np--;
      } else {
        ip++;
      }
    }
  }
  return nc;
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_mesh_cand(RebuildArgs a) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  unsigned long long nc = 0;
  if (i < a.n) nc = (unsigned long long)mesh_cand_one<T>(a, i);
  // candidate total, one atomic per wavefront
  for (int off = 32; off > 0; off >>= 1) nc += __shfl_xor(nc, off, 64);
  if ((threadIdx.x & 63) == 0 && nc) atomicAdd(&a.ctrl->nmesh_cand, nc);
}

// the first step of a run: meshes rotated in place
__global__ __launch_bounds__(kBlock) void k_mesh_move(MeshMoveArgs a) {
  if (a.guard && a.step >= a.ctrl->stop_at) return;
  mesh_move_block(a, blockIdx.x * kBlock + threadIdx.x, a.step);
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static inline int nblk(int n, int b) { return (n + b - 1) / b; }

int scan_tmp_size(int n) { return nblk(n, kScanPer) + 16; }

hipError_t launch_rebuild(const RebuildArgs &a, hipStream_t s) {
  const int nb = nblk(a.n, kBlock);
  const int ncb = nblk(a.ncell, kBlock);
  hipLaunchKernelGGL(k_reset_build_counters, dim3(1), dim3(64), 0, s, a.ctrl);
  hipLaunchKernelGGL(k_zero, dim3(ncb), dim3(kBlock), 0, s, a.cell_count, a.ncell);
  if (a.n > 0) hipLaunchKernelGGL(k_wrap_keys, dim3(nb), dim3(kBlock), 0, s, a);
  const int nsb = nblk(a.ncell, kScanPer);
  hipLaunchKernelGGL(k_scan_block, dim3(nsb), dim3(256), 0, s, a.cell_count, a.cell_start,
                     a.scan_tmp, a.ncell);
  hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(256), 0, s, a.scan_tmp, nsb,
                     a.cell_start + a.ncell);
  hipLaunchKernelGGL(k_scan_add, dim3(nsb), dim3(256), 0, s, a.cell_start, a.scan_tmp, a.ncell);
  hipLaunchKernelGGL(k_zero, dim3(ncb), dim3(kBlock), 0, s, a.cell_count, a.ncell);
  if (a.n > 0) {
    hipLaunchKernelGGL(k_scatter, dim3(nb), dim3(kBlock), 0, s, a);
    hipLaunchKernelGGL(k_cell_sort, dim3(ncb), dim3(kBlock), 0, s, a);
    if (a.mixed) {
      hipLaunchKernelGGL(k_gather<float>, dim3(nb), dim3(kBlock), 0, s, a);
      hipLaunchKernelGGL(k_build<float>, dim3(nb), dim3(kBlock), 0, s, a);
    } else {
      hipLaunchKernelGGL(k_gather<double>, dim3(nb), dim3(kBlock), 0, s, a);
      hipLaunchKernelGGL(k_build<double>, dim3(nb), dim3(kBlock), 0, s, a);
    }
  }
  return hipGetLastError();
}

template <typename T, bool MESH>
static void launch_step_variant(const StepArgs &a, int nb, hipStream_t s) {
  const int B = kStepVariantBlock[a.variant];
  if (a.prm.generic) {
    // the sibling contact models run in the default tile shape or its
    // wave-compacted form
    if (a.variant == 5)
      hipLaunchKernelGGL((k_step<T, MESH, 256, 4, true, 1>), dim3(nb), dim3(B), 0, s, a);
    else if (a.variant == 7)
      hipLaunchKernelGGL((k_step<T, MESH, 256, 4, true, 2>), dim3(nb), dim3(B), 0, s, a);
    else
      hipLaunchKernelGGL((k_step<T, MESH, 256, 4, true>), dim3(nb), dim3(B), 0, s, a);
    return;
  }
  switch (a.variant) {
    case 1:
      hipLaunchKernelGGL((k_step<T, MESH, 256, 2>), dim3(nb), dim3(B), 0, s, a);
      break;
    case 2:
      hipLaunchKernelGGL((k_step<T, MESH, 256, 8>), dim3(nb), dim3(B), 0, s, a);
      break;
    case 3:
      hipLaunchKernelGGL((k_step<T, MESH, 128, 4>), dim3(nb), dim3(B), 0, s, a);
      break;
    case 4:
      hipLaunchKernelGGL((k_step<T, MESH, 64, 4>), dim3(nb), dim3(B), 0, s, a);
      break;
    case 5:
      hipLaunchKernelGGL((k_step<T, MESH, 256, 4, false, 1>), dim3(nb), dim3(B), 0, s, a);
      break;
    case 6:
      hipLaunchKernelGGL((k_step<T, MESH, 256, 8, false, 1>), dim3(nb), dim3(B), 0, s, a);
      break;
    case 7:
      hipLaunchKernelGGL((k_step<T, MESH, 256, 4, false, 2>), dim3(nb), dim3(B), 0, s, a);
      break;
    case 8:
      hipLaunchKernelGGL((k_step<T, MESH, 256, 8, false, 2>), dim3(nb), dim3(B), 0, s, a);
      break;
    case 9:
      hipLaunchKernelGGL((k_step<T, MESH, 256, 4, false, 3>), dim3(nb), dim3(B), 0, s, a);
      break;
    case 10:
      hipLaunchKernelGGL((k_step<T, MESH, 128, 4, false, 3>), dim3(nb), dim3(B), 0, s, a);
      break;
    default:
      hipLaunchKernelGGL((k_step<T, MESH, 256, 4>), dim3(nb), dim3(B), 0, s, a);
      break;
  }
}

hipError_t launch_step(const StepArgs &a, hipStream_t s) {
  if (a.variant < 0 || a.variant >= kNumStepVariants) return hipErrorInvalidValue;
  if (a.prm.generic && a.variant != 0 && a.variant != 5 && a.variant != 7) return hipErrorInvalidValue;
  const bool mesh = a.prm.nmesh > 0;
  // pblocks (tiles walked by the particle blocks) is set by the caller;
  // trailing blocks move the meshes when asked to
  const int B = kStepVariantBlock[a.variant];
  const int nb = a.pblocks + ((mesh && a.move_mesh) ? nblk(a.mv.ntri, B) : 0);
  if (a.mixed) {
    if (mesh)
      launch_step_variant<float, true>(a, nb, s);
    else
      launch_step_variant<float, false>(a, nb, s);
  } else {
    if (mesh)
      launch_step_variant<double, true>(a, nb, s);
    else
      launch_step_variant<double, false>(a, nb, s);
  }
  return hipGetLastError();
}

hipError_t launch_pair_local_count(const StepArgs &a, const PairLocalArgs &o, hipStream_t s) {
  if (a.n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_pair_local_count, dim3(nblk(a.n, kBlock)), dim3(kBlock), 0, s, a, o);
  return hipGetLastError();
}

hipError_t launch_pair_local_fill(const StepArgs &a, const PairLocalArgs &o, hipStream_t s) {
  if (a.n <= 0) return hipSuccess;
  const dim3 g(nblk(a.n, kBlock)), b(kBlock);
  if (a.mixed) {
    if (a.prm.generic)
      hipLaunchKernelGGL((k_pair_local_fill<float, true>), g, b, 0, s, a, o);
    else
      hipLaunchKernelGGL((k_pair_local_fill<float, false>), g, b, 0, s, a, o);
  } else {
    if (a.prm.generic)
      hipLaunchKernelGGL((k_pair_local_fill<double, true>), g, b, 0, s, a, o);
    else
      hipLaunchKernelGGL((k_pair_local_fill<double, false>), g, b, 0, s, a, o);
  }
  return hipGetLastError();
}

template <bool FILL>
static hipError_t launch_wall_local(const StepArgs &a, const WallLocalArgs &o, hipStream_t s) {
  if (a.n <= 0) return hipSuccess;
  const dim3 g(nblk(a.n, kBlock)), b(kBlock);
  if (a.mixed) {
    if (a.prm.generic)
      hipLaunchKernelGGL((FILL ? k_wall_local_fill<float, true> : k_wall_local_count<float, true>), g, b, 0, s, a, o);
    else
      hipLaunchKernelGGL((FILL ? k_wall_local_fill<float, false> : k_wall_local_count<float, false>), g, b, 0, s, a,
                         o);
  } else {
    if (a.prm.generic)
      hipLaunchKernelGGL((FILL ? k_wall_local_fill<double, true> : k_wall_local_count<double, true>), g, b, 0, s, a,
                         o);
    else
      hipLaunchKernelGGL((FILL ? k_wall_local_fill<double, false> : k_wall_local_count<double, false>), g, b, 0, s,
                         a, o);
  }
  return hipGetLastError();
}

hipError_t launch_wall_local_count(const StepArgs &a, const WallLocalArgs &o, hipStream_t s) {
  return launch_wall_local<false>(a, o, s);
}

hipError_t launch_wall_local_fill(const StepArgs &a, const WallLocalArgs &o, hipStream_t s) {
  return launch_wall_local<true>(a, o, s);
}

hipError_t launch_initial(const StepArgs &a, hipStream_t s) {
  if (a.n <= 0) return hipSuccess;
  const int nb = nblk(a.n, kBlock);
  if (a.mixed)
    hipLaunchKernelGGL(k_initial<float>, dim3(nb), dim3(kBlock), 0, s, a);
  else
    hipLaunchKernelGGL(k_initial<double>, dim3(nb), dim3(kBlock), 0, s, a);
  return hipGetLastError();
}

static void launch_scan(const int *in, int *out, int *tmp, int n, hipStream_t s) {
  const int nsb = nblk(n, kScanPer);
  hipLaunchKernelGGL(k_scan_block, dim3(nsb), dim3(256), 0, s, in, out, tmp, n);
  hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(256), 0, s, tmp, nsb, out + n);
  hipLaunchKernelGGL(k_scan_add, dim3(nsb), dim3(256), 0, s, out, tmp, n);
}

hipError_t launch_exclusive_scan(const int *in, int *out, int *tmp, int n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  launch_scan(in, out, tmp, n, s);
  return hipGetLastError();
}

hipError_t launch_mesh_bin_count(const RebuildArgs &a, int *total, hipStream_t s) {
  if (a.ntri <= 0) return hipSuccess;
  const int nt = nblk(a.ntri, kBlock);
  hipLaunchKernelGGL(k_mesh_geom, dim3(nt), dim3(kBlock), 0, s, a);
  hipLaunchKernelGGL(k_tri_count, dim3(nblk(a.ntri, kWavesPerBlock)), dim3(kBlock), 0, s, a);
  // toff has ntri+1 entries; the scan writes the total at toff[ntri]
  launch_scan(a.tcount, a.toff, a.scan_tmp, a.ntri, s);
  const hipError_t e = hipMemcpyAsync(total, a.toff + a.ntri, sizeof(int),
                                      hipMemcpyDeviceToDevice, s);
  if (e != hipSuccess) return e;
  return hipGetLastError();
}

hipError_t launch_mesh_bin_fill(const RebuildArgs &a, hipStream_t s) {
  if (a.ntri <= 0) return hipSuccess;
  const int ncb = nblk(a.ncell, kBlock);
  hipLaunchKernelGGL(k_tri_fill, dim3(nblk(a.ntri, kWavesPerBlock)), dim3(kBlock), 0, s, a);
  const int np = a.tpair_cap;  // exactly the pair total (sized by the caller)
  hipLaunchKernelGGL(k_zero, dim3(ncb), dim3(kBlock), 0, s, a.tcell_count, a.ncell);
  const int npi = np * a.prm.ninst;
  if (np > 0) hipLaunchKernelGGL(k_tcell_hist, dim3(nblk(npi, kBlock)), dim3(kBlock), 0, s, a, np);
  launch_scan(a.tcell_count, a.tcell_start, a.scan_tmp, a.ncell, s);
  hipLaunchKernelGGL(k_zero, dim3(ncb), dim3(kBlock), 0, s, a.tcell_count, a.ncell);
  if (np > 0)
    hipLaunchKernelGGL(k_tcell_scatter, dim3(nblk(npi, kBlock)), dim3(kBlock), 0, s, a, np);
  if (a.n > 0) {
    if (a.mixed)
      hipLaunchKernelGGL(k_mesh_cand<float>, dim3(nblk(a.n, kBlock)), dim3(kBlock), 0, s, a);
    else
      hipLaunchKernelGGL(k_mesh_cand<double>, dim3(nblk(a.n, kBlock)), dim3(kBlock), 0, s, a);
  }
  return hipGetLastError();
}

hipError_t launch_mesh_move(const MeshMoveArgs &a, hipStream_t s) {
  if (a.ntri <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_mesh_move, dim3(nblk(a.ntri, kBlock)), dim3(kBlock), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_set_base(Ctrl *ctrl, long long step, long long ago, hipStream_t s) {
  hipLaunchKernelGGL(k_set_base, dim3(1), dim3(64), 0, s, ctrl, step, ago);
  return hipGetLastError();
}

hipError_t launch_set_stop(Ctrl *ctrl, long long v, hipStream_t s, unsigned long long *inst_stats,
                           int nstat) {
  const int nb = inst_stats ? nblk(nstat, 256) : 1;
  hipLaunchKernelGGL(k_set_stop, dim3(nb), dim3(256), 0, s, ctrl, v, inst_stats, nstat);
  return hipGetLastError();
}

}  // namespace dem
