/*
 * CPU oracle, test infrastructure only: the superquadric DEM step in plain C,
 * FP64, one thread. Nothing in the product links or calls this file.
 *
 * Follows:
 *   Verlet::setup / Verlet::run step order (verlet.cpp:134-205, 264-392)
 *   FixNVEAsphereBase::initial_integrate / final_integrate, integration_scheme 1
 *       (the default no-squish splitting; fix_nve_asphere_base.cpp:355-460,
 *       529-610) with calc_conjqm / no_squish_rotate
 *       (math_extra_liggghts_nonspherical.cpp:175-225) and MathExtra::quatvec /
 *       invquatvec / qnormalize (math_extra.h:546-616)
 *   Neighbor::decide / check_distance (neighbor.cpp:1357-1461), the granular
 *       half list predicate on the bounding radii (neigh_gran.cpp:571-582) and
 *       the history carried over a rebuild for pairs flagged as touching
 *       (fix_contact_history.cpp:305-425, neigh_gran.cpp:590-615)
 *   PairGran compute loop (pair_gran_base.h:255-430) with the superquadric
 *       surface model (surface_model_superquadric.h:108-220: checkSurfaceIntersect,
 *       then surfacesIntersectNonSpherical, math_extra_liggghts_nonspherical.h:
 *       536-577), NormalModel<HERTZ> (normal_model_hertz.h:202-405, contact-point
 *       torque of the normal force) and TangentialModel<TANGENTIAL_HISTORY>
 *       (tangential_model_history.h:130-369, contact-point lever arms)
 *   FixWallGran::post_force_primitive for superquadrics (fix_wall_gran.cpp:
 *       1024-1153) with Superquadric::plane_intersection and
 *       point_wall_projection (math_extra_liggghts_nonspherical.cpp:48-66), and
 *       the wall variant of the models (fix_wall_gran_base.h:148-361)
 *   FixGravity::post_force (fix_gravity.cpp:301-350)
 *   AtomVecSuperquadric::data_atom (atom_vec_superquadric.cpp:1150-1212):
 *       bounding radius, volume, mass, principal inertia
 *
 * Two places differ from the reference on purpose:
 *   - post_force_primitive leaves the whole wall loop (`break`,
 *     fix_wall_gran.cpp:1118-1124) for a superquadric whose bounding sphere
 *     touches the plane but whose shape does not; since deltan is minus a
 *     distance there, the branch never runs for superquadrics and the oracle
 *     has none.
 *   - the contact point kept in the pair history is a global point; when a
 *     periodic wrap moves particle i at a rebuild, the carried point moves by
 *     the same shift (the reference carries it unshifted into the new image).
 */
#include "sqdem_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "sq_oracle.h"

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

static void *xcalloc(size_t n, size_t sz) {
  void *p = calloc(n ? n : 1, sz);
  if (!p) abort();
  return p;
}

static double beta_fn(double a, double b) { return exp(lgamma(a) + lgamma(b) - lgamma(a + b)); }

void orc_sqs_mass_props(const double *shape, const double *blk, double dens, double *volume, double *I) {
  const double eps1 = 2.0 / blk[0], eps2 = 2.0 / blk[1];
  const double a1 = shape[0], a2 = shape[1], a3 = shape[2];
  *volume = 2.0 * a1 * a2 * a3 * eps1 * eps2 * beta_fn(0.5 * eps1, eps1 + 1.0) * beta_fn(0.5 * eps2, 0.5 * eps2 + 1.0);
  const double bA = beta_fn(1.5 * eps2, 0.5 * eps2) * beta_fn(0.5 * eps1, 2.0 * eps1 + 1.0);
  const double bB = beta_fn(0.5 * eps2, 0.5 * eps2 + 1.0) * beta_fn(1.5 * eps1, eps1 + 1.0);
  I[0] = 0.5 * a1 * a2 * a3 * eps1 * eps2 * (a2 * a2 * bA + 4.0 * a3 * a3 * bB) * dens;
  I[1] = 0.5 * a1 * a2 * a3 * eps1 * eps2 * (a1 * a1 * bA + 4.0 * a3 * a3 * bB) * dens;
  I[2] = 0.5 * a1 * a2 * a3 * eps1 * eps2 * (a1 * a1 + a2 * a2) * bA * dens;
}

struct orc_sqs {
  orc_sqs_params p;
  orc_materials m;
  int n;
  double *x, *v, *w, *L, *q, *shape, *blk, *mass, *rv, *rb, *I, *f, *t, *xhold, *wrap;
  int *type;
  int64_t np, cap;
  int *pi, *pj, *flag;
  double *hist;
  double *whist;
  int64_t ntimestep, nrebuild, newton, touch_pair, touch_wall;
  double cutneighmax;
};

/* MathExtraLiggghtsNonspherical::quat_to_mat */
static void rot_of(const double *q, double *m) {
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

static void mv(const double *m, const double *v, double *r) {
  r[0] = m[0] * v[0] + m[1] * v[1] + m[2] * v[2];
  r[1] = m[3] * v[0] + m[4] * v[1] + m[5] * v[2];
  r[2] = m[6] * v[0] + m[7] * v[1] + m[8] * v[2];
}

static void tmv(const double *m, const double *v, double *r) {
  r[0] = m[0] * v[0] + m[3] * v[1] + m[6] * v[2];
  r[1] = m[1] * v[0] + m[4] * v[1] + m[7] * v[2];
  r[2] = m[2] * v[0] + m[5] * v[1] + m[8] * v[2];
}

static void cross(const double *a, const double *b, double *r) {
  r[0] = a[1] * b[2] - a[2] * b[1];
  r[1] = a[2] * b[0] - a[0] * b[2];
  r[2] = a[0] * b[1] - a[1] * b[0];
}

/* MathExtra::quatvec, invquatvec, qnormalize */
static void quatvec(const double *a, const double *b, double *c) {
  c[0] = -a[1] * b[0] - a[2] * b[1] - a[3] * b[2];
  c[1] = a[0] * b[0] + a[2] * b[2] - a[3] * b[1];
  c[2] = a[0] * b[1] + a[3] * b[0] - a[1] * b[2];
  c[3] = a[0] * b[2] + a[1] * b[1] - a[2] * b[0];
}

static void invquatvec(const double *a, const double *b, double *c) {
  c[0] = -a[1] * b[0] + a[0] * b[1] + a[3] * b[2] - a[2] * b[3];
  c[1] = -a[2] * b[0] - a[3] * b[1] + a[0] * b[2] + a[1] * b[3];
  c[2] = -a[3] * b[0] + a[2] * b[1] - a[1] * b[2] + a[0] * b[3];
}

static void qnormalize(double *q) {
  const double norm = 1.0 / sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
  q[0] *= norm;
  q[1] *= norm;
  q[2] *= norm;
  q[3] *= norm;
}

/* no_squish_rotate (math_extra_liggghts_nonspherical.cpp:175-216) */
static void no_squish_rotate(int k, double *p, double *q, const double *inertia, double dt) {
  double kp[4] = {0, 0, 0, 0}, kq[4] = {0, 0, 0, 0};
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

orc_sqs *orc_sqs_create(const orc_sqs_params *p, const orc_materials *m, int n, const double *x,
                        const double *v, const double *omega, const double *quat, const double *shape,
                        const double *blk, const double *density, const int *type) {
  orc_sqs *s = (orc_sqs *)xcalloc(1, sizeof(orc_sqs));
  s->p = *p;
  s->m = *m;
  s->n = n;
  const size_t N = (size_t)n;
  s->x = (double *)xcalloc(3 * N, sizeof(double));
  s->v = (double *)xcalloc(3 * N, sizeof(double));
  s->w = (double *)xcalloc(3 * N, sizeof(double));
  s->L = (double *)xcalloc(3 * N, sizeof(double));
  s->q = (double *)xcalloc(4 * N, sizeof(double));
  s->shape = (double *)xcalloc(3 * N, sizeof(double));
  s->blk = (double *)xcalloc(2 * N, sizeof(double));
  s->mass = (double *)xcalloc(N, sizeof(double));
  s->rv = (double *)xcalloc(N, sizeof(double));
  s->rb = (double *)xcalloc(N, sizeof(double));
  s->I = (double *)xcalloc(3 * N, sizeof(double));
  s->f = (double *)xcalloc(3 * N, sizeof(double));
  s->t = (double *)xcalloc(3 * N, sizeof(double));
  s->xhold = (double *)xcalloc(3 * N, sizeof(double));
  s->wrap = (double *)xcalloc(3 * N, sizeof(double));
  s->type = (int *)xcalloc(N, sizeof(int));
  s->whist = (double *)xcalloc(N * (size_t)(p->nplanes > 0 ? p->nplanes : 1) * 3, sizeof(double));
  memcpy(s->x, x, 3 * N * sizeof(double));
  memcpy(s->v, v, 3 * N * sizeof(double));
  memcpy(s->w, omega, 3 * N * sizeof(double));
  memcpy(s->q, quat, 4 * N * sizeof(double));
  memcpy(s->shape, shape, 3 * N * sizeof(double));
  memcpy(s->blk, blk, 2 * N * sizeof(double));
  memcpy(s->type, type, N * sizeof(int));
  double rbmax = 0.0;
  for (int i = 0; i < n; i++) {
    double vol;
    orc_sqs_mass_props(&shape[3 * i], &blk[2 * i], density[i], &vol, &s->I[3 * i]);
    s->mass[i] = vol * density[i];
    s->rv[i] = cbrt(0.75 * vol / M_PI);
    s->rb[i] = orc_sq_bounding_radius(&shape[3 * i], &blk[2 * i]);
    if (s->rb[i] > rbmax) rbmax = s->rb[i];
    /* angular momentum of the given angular velocity: R diag(I) R^T omega */
    double R[9], wb[3], mb[3];
    rot_of(&s->q[4 * i], R);
    tmv(R, &s->w[3 * i], wb);
    for (int k = 0; k < 3; k++) mb[k] = s->I[3 * i + k] * wb[k];
    mv(R, mb, &s->L[3 * i]);
  }
  s->cutneighmax = 2.0 * rbmax + p->skin;
  return s;
}

void orc_sqs_destroy(orc_sqs *s) {
  if (!s) return;
  free(s->x); free(s->v); free(s->w); free(s->L); free(s->q); free(s->shape); free(s->blk);
  free(s->mass); free(s->rv); free(s->rb); free(s->I); free(s->f); free(s->t); free(s->xhold);
  free(s->wrap); free(s->type); free(s->whist); free(s->pi); free(s->pj); free(s->flag); free(s->hist);
  free(s);
}

/* periodic wrap (Domain::pbc on rebuild steps), remembering each shift */
static void wrap_all(orc_sqs *s) {
  for (int i = 0; i < s->n; i++)
    for (int d = 0; d < 3; d++) {
      s->wrap[3 * i + d] = 0.0;
      if (!s->p.periodic[d]) continue;
      const double lo = s->p.boxlo[d], hi = s->p.boxhi[d], L = hi - lo;
      double *xi = &s->x[3 * i + d];
      const double x0 = *xi;
      /* Domain::pbc, domain.cpp:530-620 (orthogonal box) */
      if (*xi < lo) *xi += L;
      if (*xi >= hi) {
        *xi -= L;
        if (*xi < lo) *xi = lo;
      }
      s->wrap[3 * i + d] = *xi - x0;
    }
}

/* the periodic image of j nearest to i: x_j + k L, k in {-1, 0, 1} */
static void image_of(const orc_sqs *s, int i, int j, double *xj, double *del) {
  for (int d = 0; d < 3; d++) {
    xj[d] = s->x[3 * j + d];
    if (s->p.periodic[d]) {
      const double L = s->p.boxhi[d] - s->p.boxlo[d];
      const double dd = s->x[3 * i + d] - xj[d];
      if (dd > 0.5 * L)
        xj[d] = xj[d] + L;
      else if (dd < -0.5 * L)
        xj[d] = xj[d] - L;
    }
    del[d] = s->x[3 * i + d] - xj[d];
  }
}

/* half list on the bounding radii: rsq <= (rb_i + rb_j + skin)^2, with the
 * history of pairs flagged as touching carried over when their bounding
 * spheres overlap (the contact point shifted with i's wrap) */
static void build(orc_sqs *s) {
  const int n = s->n;
  const int64_t onp = s->np;
  int *opi = s->pi, *opj = s->pj, *ofl = s->flag;
  double *oh = s->hist;
  /* first index of each i in the old list (old list is i-major, j ascending) */
  int64_t *ofirst = (int64_t *)xcalloc((size_t)n + 1, sizeof(int64_t));
  for (int64_t e = 0; e < onp; e++) ofirst[opi[e] + 1]++;
  for (int i = 0; i < n; i++) ofirst[i + 1] += ofirst[i];
  int64_t cnt = 0;
  for (int pass = 0; pass < 2; pass++) {
    int64_t np = 0;
    for (int i = 0; i < n; i++)
      for (int j = i + 1; j < n; j++) {
        double xj[3], del[3];
        image_of(s, i, j, xj, del);
        const double rsq = del[0] * del[0] + del[1] * del[1] + del[2] * del[2];
        const double cut = s->rb[i] + s->rb[j] + s->p.skin;
        if (rsq > cut * cut) continue;
        if (pass == 1) {
          s->pi[np] = i;
          s->pj[np] = j;
          double *h = &s->hist[ORC_SQS_HIST * np];
          memset(h, 0, sizeof(double) * ORC_SQS_HIST);
          int fl = 0;
          const double radsum = s->rb[i] + s->rb[j];
          if (rsq < radsum * radsum)
            for (int64_t e = ofirst[i]; e < ofirst[i + 1]; e++)
              if (opj[e] == j && ofl[e]) {
                memcpy(h, &oh[ORC_SQS_HIST * e], sizeof(double) * ORC_SQS_HIST);
                for (int d = 0; d < 3; d++) h[2 + d] += s->wrap[3 * i + d];
                fl = 1;
                break;
              }
          s->flag[np] = fl;
        }
        np++;
      }
    if (pass == 0) {
      cnt = np;
      s->pi = (int *)xcalloc((size_t)cnt, sizeof(int));
      s->pj = (int *)xcalloc((size_t)cnt, sizeof(int));
      s->flag = (int *)xcalloc((size_t)cnt, sizeof(int));
      s->hist = (double *)xcalloc((size_t)cnt * ORC_SQS_HIST, sizeof(double));
    }
  }
  s->np = cnt;
  free(opi);
  free(opj);
  free(ofl);
  free(oh);
  free(ofirst);
  memcpy(s->xhold, s->x, sizeof(double) * 3 * (size_t)n);
  s->nrebuild++;
}

/* Hertz normal force and Coulomb-capped tangential spring of one superquadric
 * contact, with the lever arms xci (and xcj for a pair) from the centres to
 * the contact point. is_wall: j is the wall (v_j = 0, omega_j unused). */
static void contact_force(const orc_materials *mat, int tdamp, int limit, double dt, int itype, int jtype,
                          int is_wall, double reff, double meff, double deltan, const double *en,
                          const double *vi, const double *vj, const double *wi, const double *wj,
                          const double *xci, const double *xcj, int shearupdate, double *shear, double *dFi,
                          double *dTi, double *dTj) {
  /* surfacesIntersectNonSpherical (math_extra_liggghts_nonspherical.h:536-577) */
  double vri[3], vrr[3], vrel[3];
  cross(wi, xci, vri);
  if (is_wall) {
    memcpy(vrr, vri, sizeof vri);
  } else {
    double vrj[3];
    cross(wj, xcj, vrj);
    for (int k = 0; k < 3; k++) vrr[k] = vri[k] - vrj[k];
  }
  for (int k = 0; k < 3; k++) vrel[k] = vi[k] - vj[k] + vrr[k];
  const double vn = vrel[0] * en[0] + vrel[1] * en[1] + vrel[2] * en[2];
  const double vt[3] = {vrel[0] - vn * en[0], vrel[1] - vn * en[1], vrel[2] - vn * en[2]};
  /* NormalModel<HERTZ>::surfacesIntersect (normal_model_hertz.h:202-405) */
  const double Y = mat->Yeff[itype][jtype], G = mat->Geff[itype][jtype], B = mat->betaeff[itype][jtype];
  const double sqrtval = sqrt(reff * deltan);
  const double Sn = 2. * Y * sqrtval;
  const double St = 8. * G * sqrtval;
  const double kn = 4. / 3. * Y * sqrtval;
  const double kt = St;
  const double s56 = 0.91287092917527685576161630466800355658790782499663875;
  const double gamman = -2. * s56 * B * sqrt(Sn * meff);
  const double gammat = tdamp ? -2. * s56 * B * sqrt(St * meff) : 0.0;
  const double Fn_damping = -gamman * vn;
  const double Fn_contact = kn * deltan;
  double Fn = Fn_damping + Fn_contact;
  if (limit && Fn < 0.0) Fn = 0.0;
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
  /* TangentialModel<TANGENTIAL_HISTORY>::surfacesIntersect (tangential_model_history.h:130-352) */
  if (shearupdate) {
    shear[0] += vt[0] * dt;
    shear[1] += vt[1] * dt;
    shear[2] += vt[2] * dt;
    const double rsht = shear[0] * en[0] + shear[1] * en[1] + shear[2] * en[2];
    shear[0] -= rsht * en[0];
    shear[1] -= rsht * en[1];
    shear[2] -= rsht * en[2];
  }
  const double shrmag = sqrt(shear[0] * shear[0] + shear[1] * shear[1] + shear[2] * shear[2]);
  const double xmu = mat->mu[itype][jtype];
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

int orc_sqs_contact_pair(const orc_materials *mat, int tangential_damping, int limit_force, double dt,
                         int itype, int jtype, const orc_sq_pose *a, const orc_sq_pose *b, const double *vi,
                         const double *vj, const double *wi, const double *wj, double mi, double mj, double rvi,
                         double rvj, int shearupdate, double *hist, double *dFi, double *dTi, double *dTj) {
  double vel[12], ms[2] = {mi, mj};
  memcpy(vel, vi, 3 * sizeof(double));
  memcpy(vel + 3, vj, 3 * sizeof(double));
  memcpy(vel + 6, wi, 3 * sizeof(double));
  memcpy(vel + 9, wj, 3 * sizeof(double));
  double en[3], dn, cpa[3], cpb[3], d[3];
  int it = 0;
  for (int k = 0; k < 3; k++) dFi[k] = dTi[k] = dTj[k] = 0.0;
  if (!orc_sq_check(a, b, hist, vel, ms, dt, en, &dn, cpa, cpb, d, &it)) {
    hist[7] = hist[8] = hist[9] = 0.0;
    return 0;
  }
  const double *cp = hist + 2;
  const double xci[3] = {cp[0] - a->center[0], cp[1] - a->center[1], cp[2] - a->center[2]};
  const double xcj[3] = {cp[0] - b->center[0], cp[1] - b->center[1], cp[2] - b->center[2]};
  contact_force(mat, tangential_damping, limit_force, dt, itype, jtype, 0, rvi * rvj / (rvi + rvj),
                mi * mj / (mi + mj), dn, en, vi, vj, wi, wj, xci, xcj, shearupdate, hist + 7, dFi, dTi, dTj);
  return 1;
}

static void make_pose(const orc_sqs *s, int i, const double *center, orc_sq_pose *p) {
  memcpy(p->center, center, sizeof p->center);
  memcpy(p->quat, &s->q[4 * i], sizeof p->quat);
  memcpy(p->shape, &s->shape[3 * i], sizeof p->shape);
  memcpy(p->blockiness, &s->blk[2 * i], sizeof p->blockiness);
}

static void pair_compute(orc_sqs *s, int shearupdate) {
  for (int64_t e = 0; e < s->np; e++) {
    const int i = s->pi[e], j = s->pj[e];
    double xj[3], del[3];
    image_of(s, i, j, xj, del);
    const double rsq = del[0] * del[0] + del[1] * del[1] + del[2] * del[2];
    const double radsum = s->rb[i] + s->rb[j];
    double *h = &s->hist[ORC_SQS_HIST * e];
    if (!(rsq < radsum * radsum)) continue;
    orc_sq_pose A, B;
    make_pose(s, i, &s->x[3 * i], &A);
    make_pose(s, j, xj, &B);
    double vel[12], ms[2] = {s->mass[i], s->mass[j]};
    memcpy(vel, &s->v[3 * i], 3 * sizeof(double));
    memcpy(vel + 3, &s->v[3 * j], 3 * sizeof(double));
    memcpy(vel + 6, &s->w[3 * i], 3 * sizeof(double));
    memcpy(vel + 9, &s->w[3 * j], 3 * sizeof(double));
    double en[3], dn, cpa[3], cpb[3], d[3];
    int it = 0;
    const int c = orc_sq_check(&A, &B, h, vel, ms, s->p.dt, en, &dn, cpa, cpb, d, &it);
    s->newton += it;
    if (!c) {
      /* surfacesClose: the tangential model zeroes the shear (tangential_model_history.h:358-369) */
      h[7] = h[8] = h[9] = 0.0;
      continue;
    }
    s->flag[e] = 1;
    s->touch_pair++;
    const double reff = s->rv[i] * s->rv[j] / (s->rv[i] + s->rv[j]);
    const double meff = s->mass[i] * s->mass[j] / (s->mass[i] + s->mass[j]);
    const double *cp = h + 2;
    const double xci[3] = {cp[0] - s->x[3 * i], cp[1] - s->x[3 * i + 1], cp[2] - s->x[3 * i + 2]};
    const double xcj[3] = {cp[0] - xj[0], cp[1] - xj[1], cp[2] - xj[2]};
    double dFi[3], dTi[3], dTj[3];
    contact_force(&s->m, s->p.tangential_damping, s->p.limit_force, s->p.dt, s->type[i], s->type[j], 0, reff,
                  meff, dn, en, &s->v[3 * i], &s->v[3 * j], &s->w[3 * i], &s->w[3 * j], xci, xcj, shearupdate,
                  h + 7, dFi, dTi, dTj);
    for (int k = 0; k < 3; k++) {
      s->f[3 * i + k] += dFi[k];
      s->t[3 * i + k] += dTi[k];
      s->f[3 * j + k] += -dFi[k];
      s->t[3 * j + k] += dTj[k];
    }
  }
}

static void wall_compute(orc_sqs *s, int shearupdate) {
  const double vwall[3] = {0.0, 0.0, 0.0};
  for (int k = 0; k < s->p.nplanes; k++) {
    const int ax = s->p.plane_axis[k];
    const double pos = s->p.plane_pos[k];
    for (int i = 0; i < s->n; i++) {
      double *sh = &s->whist[3 * ((size_t)i * s->p.nplanes + k)];
      const double *xi = &s->x[3 * i];
      /* Plane<dim>::resolveContact (primitive_wall_definitions.h:133-142) on the bounding radius */
      double delta[3] = {0.0, 0.0, 0.0}, deltan;
      delta[ax] = pos - xi[ax];
      if (xi[ax] > pos)
        deltan = xi[ax] - pos - s->rb[i];
      else
        deltan = pos - xi[ax] - s->rb[i];
      if (deltan > 0.0) {
        sh[0] = sh[1] = sh[2] = 0.0;
        continue;
      }
      const double scp[3] = {xi[0] + delta[0], xi[1] + delta[1], xi[2] + delta[2]};
      orc_sq_pose P;
      make_pose(s, i, xi, &P);
      double cp[3], low[3];
      const int hit = orc_sq_plane(&P, delta, scp, cp, low);
      /* point_wall_projection (math_extra_liggghts_nonspherical.cpp:48-66) */
      const double A = delta[0], Bc = delta[1], C = delta[2];
      const double dist = fabs((A * (cp[0] - scp[0]) + Bc * (cp[1] - scp[1]) + C * (cp[2] - scp[2])) /
                               sqrt(A * A + Bc * Bc + C * C));
      if (!hit) {
        sh[0] = sh[1] = sh[2] = 0.0;
        continue;
      }
      s->touch_wall++;
      const double sd[3] = {-delta[0], -delta[1], -delta[2]};
      const double dinv = 1.0 / sqrt(sd[0] * sd[0] + sd[1] * sd[1] + sd[2] * sd[2]);
      const double en[3] = {sd[0] * dinv, sd[1] * dinv, sd[2] * dinv};
      const double xci[3] = {cp[0] - xi[0], cp[1] - xi[1], cp[2] - xi[2]};
      double dFi[3], dTi[3];
      contact_force(&s->m, s->p.tangential_damping, s->p.limit_force, s->p.dt, s->type[i], s->p.plane_type[k], 1,
                    s->rv[i], s->mass[i], dist, en, &s->v[3 * i], vwall, &s->w[3 * i], NULL, xci, NULL,
                    shearupdate, sh, dFi, dTi, NULL);
      for (int d = 0; d < 3; d++) {
        s->f[3 * i + d] += dFi[d];
        s->t[3 * i + d] += dTi[d];
      }
    }
  }
}

static void gravity_compute(orc_sqs *s) {
  if (!s->p.has_gravity) return;
  for (int i = 0; i < s->n; i++)
    for (int d = 0; d < 3; d++) s->f[3 * i + d] += s->mass[i] * s->p.grav_acc[d];
}

static void forces(orc_sqs *s, int shearupdate) {
  memset(s->f, 0, sizeof(double) * 3 * (size_t)s->n);
  memset(s->t, 0, sizeof(double) * 3 * (size_t)s->n);
  s->newton = s->touch_pair = s->touch_wall = 0;
  pair_compute(s, shearupdate);
  if (s->p.gravity_first) {
    gravity_compute(s);
    wall_compute(s, shearupdate);
  } else {
    wall_compute(s, shearupdate);
    gravity_compute(s);
  }
}

void orc_sqs_setup(orc_sqs *s) {
  wrap_all(s);
  build(s);
  forces(s, 0);
}

static void initial_integrate(orc_sqs *s) {
  const double dtv = s->p.dt, dtf = 0.5 * s->p.dt, dtf2 = 2.0 * dtf, dtq = 0.5 * dtv;
  for (int i = 0; i < s->n; i++) {
    double *x = &s->x[3 * i], *v = &s->v[3 * i], *L = &s->L[3 * i], *q = &s->q[4 * i], *w = &s->w[3 * i];
    const double *f = &s->f[3 * i], *t = &s->t[3 * i], *I = &s->I[3 * i];
    double R[9], mbody[3], conjqm[4], tbody[3], fquat[4], wbody[3];
    rot_of(q, R);
    const double dtfm = dtf / s->mass[i];
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
    qnormalize(q);
    rot_of(q, R);
    invquatvec(q, conjqm, mbody);
    for (int k = 0; k < 3; k++) mbody[k] *= 0.5;
    for (int k = 0; k < 3; k++) wbody[k] = mbody[k] / I[k];
    mv(R, mbody, L);
    mv(R, wbody, w);
  }
}

static void final_integrate(orc_sqs *s) {
  const double dtf = 0.5 * s->p.dt, dtf2 = 2.0 * dtf;
  for (int i = 0; i < s->n; i++) {
    double *v = &s->v[3 * i], *L = &s->L[3 * i], *q = &s->q[4 * i], *w = &s->w[3 * i];
    const double *f = &s->f[3 * i], *t = &s->t[3 * i], *I = &s->I[3 * i];
    double R[9], mbody[3], conjqm[4], tbody[3], fquat[4], wbody[3];
    const double dtfm = dtf / s->mass[i];
    rot_of(q, R);
    for (int k = 0; k < 3; k++) v[k] += dtfm * f[k];
    tmv(R, t, tbody);
    tmv(R, L, mbody);
    quatvec(q, mbody, conjqm);
    for (int k = 0; k < 4; k++) conjqm[k] *= 2.0;
    quatvec(q, tbody, fquat);
    for (int k = 0; k < 4; k++) conjqm[k] += dtf2 * fquat[k];
    invquatvec(q, conjqm, mbody);
    for (int k = 0; k < 3; k++) mbody[k] *= 0.5;
    for (int k = 0; k < 3; k++) wbody[k] = mbody[k] / I[k];
    mv(R, mbody, L);
    mv(R, wbody, w);
  }
}

static int check_distance(const orc_sqs *s) {
  const double trigger = 0.5 * s->p.skin;
  const double tsq = trigger * trigger;
  for (int i = 0; i < s->n; i++) {
    const double dx = s->x[3 * i] - s->xhold[3 * i];
    const double dy = s->x[3 * i + 1] - s->xhold[3 * i + 1];
    const double dz = s->x[3 * i + 2] - s->xhold[3 * i + 2];
    if (dx * dx + dy * dy + dz * dz > tsq) return 1;
  }
  return 0;
}

void orc_sqs_run(orc_sqs *s, int nsteps) {
  for (int step = 0; step < nsteps; step++) {
    s->ntimestep++;
    initial_integrate(s);
    if (check_distance(s)) {
      wrap_all(s);
      build(s);
    }
    forces(s, 1);
    final_integrate(s);
  }
}

void orc_sqs_get_state(const orc_sqs *s, double *x, double *v, double *omega, double *angmom, double *quat,
                       double *f, double *torque) {
  const size_t N = (size_t)s->n;
  if (x) memcpy(x, s->x, 3 * N * sizeof(double));
  if (v) memcpy(v, s->v, 3 * N * sizeof(double));
  if (omega) memcpy(omega, s->w, 3 * N * sizeof(double));
  if (angmom) memcpy(angmom, s->L, 3 * N * sizeof(double));
  if (quat) memcpy(quat, s->q, 4 * N * sizeof(double));
  if (f) memcpy(f, s->f, 3 * N * sizeof(double));
  if (torque) memcpy(torque, s->t, 3 * N * sizeof(double));
}

void orc_sqs_get_props(const orc_sqs *s, double *mass, double *rvol, double *rbound, double *inertia) {
  const size_t N = (size_t)s->n;
  if (mass) memcpy(mass, s->mass, N * sizeof(double));
  if (rvol) memcpy(rvol, s->rv, N * sizeof(double));
  if (rbound) memcpy(rbound, s->rb, N * sizeof(double));
  if (inertia) memcpy(inertia, s->I, 3 * N * sizeof(double));
}

int64_t orc_sqs_get_pairs(const orc_sqs *s, int64_t cap, int *ij, int *flags, double *hist) {
  if (cap < s->np) return -s->np;
  for (int64_t e = 0; e < s->np; e++) {
    if (ij) {
      ij[2 * e] = s->pi[e];
      ij[2 * e + 1] = s->pj[e];
    }
    if (flags) flags[e] = s->flag[e];
    if (hist) memcpy(&hist[ORC_SQS_HIST * e], &s->hist[ORC_SQS_HIST * e], sizeof(double) * ORC_SQS_HIST);
  }
  return s->np;
}

void orc_sqs_get_wall_hist(const orc_sqs *s, double *hist) {
  memcpy(hist, s->whist, sizeof(double) * 3 * (size_t)s->n * (size_t)s->p.nplanes);
}

void orc_sqs_stats(const orc_sqs *s, int64_t *npair_touch, int64_t *nwall_touch, int64_t *newton_iters,
                   int64_t *nrebuild, int64_t *ntimestep) {
  if (npair_touch) *npair_touch = s->touch_pair;
  if (nwall_touch) *nwall_touch = s->touch_wall;
  if (newton_iters) *newton_iters = s->newton;
  if (nrebuild) *nrebuild = s->nrebuild;
  if (ntimestep) *ntimestep = s->ntimestep;
}
