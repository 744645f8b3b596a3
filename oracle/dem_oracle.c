/*
 * CPU oracle for the DEM step (TEST INFRASTRUCTURE ONLY; see dem_oracle.h).
 *
 * FP64, single-threaded, compiled with -ffp-contract=off so that every
 * product and sum rounds exactly as the reference's x86-64 build does.
 * Each function names the reference file:line (LIGGGHTS-PUBLIC 3.8.0, src/)
 * whose arithmetic it follows.
 */
#include "dem_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define MY_PI 3.14159265358979323846 /* math_const.h */

/* ------------------------------------------------------------------------
 * material mixing
 *   Yeff    global_properties.cpp:426-433
 *   Geff    global_properties.cpp:455-464
 *   betaeff global_properties.cpp:518-541 (log e / sqrt(log^2 e + pi^2))
 *   mu, mur global_properties.cpp:548-558 (per type pair, as given)
 * ---------------------------------------------------------------------- */
void orc_mix_materials(int nt, const double *Y, const double *nu, const double *e,
                       const double *mu, const double *mur, orc_materials *out) {
  memset(out, 0, sizeof(*out));
  out->ntypes = nt;
  for (int i = 1; i <= nt; i++)
    for (int j = 1; j <= nt; j++) {
      const double Yi = Y[i - 1], Yj = Y[j - 1];
      const double vi = nu[i - 1], vj = nu[j - 1];
      out->Yeff[i][j] = 1. / ((1. - pow(vi, 2.)) / Yi + (1. - pow(vj, 2.)) / Yj);
      out->Geff[i][j] = 1. / (2. * (2. - vi) * (1. + vi) / Yi + 2. * (2. - vj) * (1. + vj) / Yj);
      const double le = log(e[(i - 1) * nt + (j - 1)]);
      out->betaeff[i][j] = le / sqrt(pow(le, 2.) + pow(MY_PI, 2.));
      out->mu[i][j] = mu[(i - 1) * nt + (j - 1)];
      out->mur[i][j] = mur ? mur[(i - 1) * nt + (j - 1)] : 0.0;
      out->loge[i][j] = le; /* createCoeffRestLog, global_properties.cpp:502-520 */
    }
}

/* MODEL_PARAMS::createKn / createKt / createGamman / createGammat (and the
 * _abs variants) / createCohesionEnergyDensity / createCharacteristicVelocity
 * (global_properties.cpp:227-240, 291-294, 735-812) with coarse-graining 1:
 * the per-type-pair values as given */
void orc_set_model_coeffs(orc_materials *m, int nt, const double *kn, const double *kt,
                          const double *gamman, const double *gammat, const double *ced,
                          double charvel) {
  const double *src[5] = {kn, kt, gamman, gammat, ced};
  double(*dst[5])[ORC_MAXTYPE + 1] = {m->kn, m->kt, m->gamman, m->gammat, m->ced};
  for (int q = 0; q < 5; q++)
    for (int i = 1; i <= nt; i++)
      for (int j = 1; j <= nt; j++) dst[q][i][j] = src[q] ? src[q][(i - 1) * nt + (j - 1)] : 0.0;
  m->charvel = charvel;
}

/* ------------------------------------------------------------------------
 * shared model chain for pair and wall contacts, in the order of
 * ContactModel::surfacesIntersect (contact_models.h:228-238):
 *   SurfaceModel<DEFAULT>::surfacesIntersect   surface_model_default.h:140-205
 *   NormalModel<HERTZ>::surfacesIntersect      normal_model_hertz.h:202-405
 *     or <HOOKE>                               normal_model_hooke.h:227-436
 *     or <HERTZ_STIFFNESS>                     normal_model_hertz_stiffness.h:106-202
 *     or <HOOKE_STIFFNESS>                     normal_model_hooke_stiffness.h:194-365
 *   CohesionModel<SJKR>::surfacesIntersect     cohesion_model_sjkr.h:86-124
 *   TangentialModel<HISTORY>::surfacesIntersect tangential_model_history.h:130-356
 *     or <NO_HISTORY>                          tangential_model_no_history.h:112-242
 *   RollingModel<CDT>::surfacesIntersect       rolling_model_cdt.h:89-167
 * ---------------------------------------------------------------------- */
typedef struct {
  int is_wall;
  double radi, radj, radsum, r, rinv, meff;
  double en[3], delta[3];
  const double *vi, *vj, *wi, *wj;
  double deltan_in; /* sidata.deltan set by the wall before the chain */
  int itype, jtype;
} contact_in;

static void model_chain(const orc_model *mdl, const orc_materials *mat, const contact_in *c,
                        double dt, int shearupdate, double *shear, double *rhist, double dFi[3],
                        double dTi[3], double dTj[3]) {
  const double enx = c->en[0], eny = c->en[1], enz = c->en[2];
  /* surface model */
  const double vr1 = c->vi[0] - c->vj[0];
  const double vr2 = c->vi[1] - c->vj[1];
  const double vr3 = c->vi[2] - c->vj[2];
  const double vn = vr1 * enx + vr2 * eny + vr3 * enz;
  const double vn1 = vn * enx, vn2 = vn * eny, vn3 = vn * enz;
  const double vt1 = vr1 - vn1, vt2 = vr2 - vn2, vt3 = vr3 - vn3;
  const double deltan = c->radsum - c->r;
  const double dx = c->delta[0], dy = c->delta[1], dz = c->delta[2];
  const double rinv = c->rinv;
  double wr1, wr2, wr3, cri, crj = 0.0;
  if (c->is_wall) {
    const double cr = c->radi - 0.5 * c->deltan_in;
    wr1 = cr * c->wi[0] * rinv;
    wr2 = cr * c->wi[1] * rinv;
    wr3 = cr * c->wi[2] * rinv;
    cri = cr;
  } else {
    cri = c->radi - 0.5 * deltan;
    crj = c->radj - 0.5 * deltan;
    wr1 = (cri * c->wi[0] + crj * c->wj[0]) * rinv;
    wr2 = (cri * c->wi[1] + crj * c->wj[1]) * rinv;
    wr3 = (cri * c->wi[2] + crj * c->wj[2]) * rinv;
  }
  const double vtr1 = vt1 - (dz * wr2 - dy * wr3);
  const double vtr2 = vt2 - (dx * wr3 - dz * wr1);
  const double vtr3 = vt3 - (dy * wr1 - dx * wr2);

  /* normal model */
  const int it = c->itype, jt = c->jtype;
  const double reff = c->is_wall ? c->radi : (c->radi * c->radj / (c->radi + c->radj));
  const double meff = c->meff;
  double kn, kt, gamman, gammat;
  if (mdl->normal_model == ORC_NORMAL_HOOKE) {
    /* normal_model_hooke.h:252-286 (viscous off) */
    const double Y = mat->Yeff[it][jt], cv = mat->charvel;
    const double sqrtval = sqrt(reff);
    kn = 16. / 15. * sqrtval * Y * pow(15. * meff * cv * cv / (16. * sqrtval * Y), 0.2);
    kt = kn;
    if (mdl->kt_to_kn) kt *= 0.285714286;
    const double lsq = mat->loge[it][jt] * mat->loge[it][jt];
    gamman = sqrt(4. * meff * kn * lsq / (lsq + MY_PI * MY_PI));
    gammat = mdl->tangential_damping ? gamman : 0.0;
  } else if (mdl->normal_model == ORC_NORMAL_HERTZ_STIFFNESS) {
    /* normal_model_hertz_stiffness.h:112-123 */
    const double polyhertz = sqrt(reff * deltan);
    kn = polyhertz * mat->kn[it][jt];
    kt = polyhertz * mat->kt[it][jt];
    gamman = polyhertz * meff * mat->gamman[it][jt];
    gammat = mdl->tangential_damping ? polyhertz * meff * mat->gammat[it][jt] : 0.0;
  } else if (mdl->normal_model == ORC_NORMAL_HOOKE_STIFFNESS) {
    /* normal_model_hooke_stiffness.h:203-226 */
    kn = mat->kn[it][jt];
    kt = mat->kt[it][jt];
    if (mdl->absolute_damping) {
      gamman = mat->gamman[it][jt];
      gammat = mat->gammat[it][jt];
    } else {
      gamman = meff * mat->gamman[it][jt];
      gammat = meff * mat->gammat[it][jt];
    }
    if (!mdl->tangential_damping) gammat = 0.0;
  } else {
    /* Hertz, normal_model_hertz.h:202-260 */
    const double sqrtval = sqrt(reff * deltan);
    const double Sn = 2. * mat->Yeff[it][jt] * sqrtval;
    const double St = 8. * mat->Geff[it][jt] * sqrtval;
    kn = 4. / 3. * mat->Yeff[it][jt] * sqrtval;
    kt = St;
    const double sqrtFiveOverSix = 0.91287092917527685576161630466800355658790782499663875;
    gamman = -2. * sqrtFiveOverSix * mat->betaeff[it][jt] * sqrt(Sn * meff);
    gammat = mdl->tangential_damping
                 ? -2. * sqrtFiveOverSix * mat->betaeff[it][jt] * sqrt(St * meff)
                 : 0.0;
  }
  kn /= 1.0; /* force->nktv2p = 1 in SI (update.cpp:223-229) */
  kt /= 1.0;
  const double Fn_damping = -gamman * vn;
  const double Fn_contact = kn * deltan;
  double Fn = Fn_damping + Fn_contact;
  if (mdl->limit_force && Fn < 0.0) Fn = 0.0;
  dFi[0] = 0.0 + Fn * enx;
  dFi[1] = 0.0 + Fn * eny;
  dFi[2] = 0.0 + Fn * enz;
  dTi[0] = dTi[1] = dTi[2] = 0.0;
  dTj[0] = dTj[1] = dTj[2] = 0.0;

  /* cohesion model: SJKR, the contact-circle area times the energy density */
  if (mdl->cohesion_model == 1) {
    const double r = c->r, ri = c->radi, rj = c->radj;
    double Acont;
    if (c->is_wall)
      Acont = (ri * ri - r * r) * MY_PI;
    else
      Acont = -MY_PI / 4 * ((r - ri - rj) * (r + ri - rj) * (r - ri + rj) * (r + ri + rj)) / (r * r);
    const double Fn_coh = -mat->ced[it][jt] * Acont;
    if (mdl->tangential_reduce) Fn += Fn_coh; /* sidata.Fn seen by the tangential model */
    dFi[0] += Fn_coh * enx;
    dFi[1] += Fn_coh * eny;
    dFi[2] += Fn_coh * enz;
  }

  const double xmu = mat->mu[it][jt];
  if (mdl->tangential_model == 1) {
    /* tangential model: no history, viscous friction capped by Coulomb */
    const double vrel = sqrt(vtr1 * vtr1 + vtr2 * vtr2 + vtr3 * vtr3);
    const double Ft_friction = xmu * fabs(Fn);
    double gamma = 0.0;
    if (Ft_friction < gammat * vrel)
      gamma = Ft_friction / vrel;
    else
      gamma = gammat;
    const double Ft1 = -gamma * vtr1;
    const double Ft2 = -gamma * vtr2;
    const double Ft3 = -gamma * vtr3;
    const double tor1 = (eny * Ft3 - enz * Ft2);
    const double tor2 = (enz * Ft1 - enx * Ft3);
    const double tor3 = (enx * Ft2 - eny * Ft1);
    dFi[0] += Ft1;
    dFi[1] += Ft2;
    dFi[2] += Ft3;
    dTi[0] += -cri * tor1;
    dTi[1] += -cri * tor2;
    dTi[2] += -cri * tor3;
    if (!c->is_wall) {
      dTj[0] += -crj * tor1;
      dTj[1] += -crj * tor2;
      dTj[2] += -crj * tor3;
    }
    goto rolling;
  }

  /* tangential model: history */
  if (shearupdate) {
    shear[0] += vtr1 * dt;
    shear[1] += vtr2 * dt;
    shear[2] += vtr3 * dt;
    const double rsht = shear[0] * enx + shear[1] * eny + shear[2] * enz;
    shear[0] -= rsht * enx;
    shear[1] -= rsht * eny;
    shear[2] -= rsht * enz;
  }
  const double shrmag = sqrt(shear[0] * shear[0] + shear[1] * shear[1] + shear[2] * shear[2]);
  double Ft1 = -(kt * shear[0]);
  double Ft2 = -(kt * shear[1]);
  double Ft3 = -(kt * shear[2]);
  const double Ft_shear = kt * shrmag;
  const double Ft_friction = xmu * fabs(Fn);
  if (Ft_shear > Ft_friction) {
    if (shrmag != 0.0) {
      const double ratio = Ft_friction / Ft_shear;
      Ft1 *= ratio;
      Ft2 *= ratio;
      Ft3 *= ratio;
      if (shearupdate) {
        shear[0] = -Ft1 / kt;
        shear[1] = -Ft2 / kt;
        shear[2] = -Ft3 / kt;
      }
    } else
      Ft1 = Ft2 = Ft3 = 0.0;
  } else {
    Ft1 -= (gammat * vtr1);
    Ft2 -= (gammat * vtr2);
    Ft3 -= (gammat * vtr3);
  }
  const double tor1 = eny * Ft3 - enz * Ft2;
  const double tor2 = enz * Ft1 - enx * Ft3;
  const double tor3 = enx * Ft2 - eny * Ft1;
  dFi[0] += Ft1;
  dFi[1] += Ft2;
  dFi[2] += Ft3;
  dTi[0] += -cri * tor1;
  dTi[1] += -cri * tor2;
  dTi[2] += -cri * tor3;
  if (!c->is_wall) {
    dTj[0] += -crj * tor1;
    dTj[1] += -crj * tor2;
    dTj[2] += -crj * tor3;
  }

rolling:
  if (mdl->rolling_epsd2) {
    /* RollingModel<EPSD2>::surfacesIntersect / calcRollTorque
     * (rolling_model_epsd2.h:98-221): an elastic rolling spring kr = kt R*^2
     * on the relative rotation (less its twist about en unless torsionTorque),
     * capped at mu_r |Fn| R*; the torque is the contact's history */
    const double rmu = mat->mur[it][jt];
    double wr_tot[3];
    if (c->is_wall) {
      wr_tot[0] = wr1;
      wr_tot[1] = wr2;
      wr_tot[2] = wr3;
    } else {
      wr_tot[0] = c->wi[0] - c->wj[0];
      wr_tot[1] = c->wi[1] - c->wj[1];
      wr_tot[2] = c->wi[2] - c->wj[2];
    }
    if (!mdl->torsion_torque) {
      const double wr_dot_delta = wr_tot[0] * enx + wr_tot[1] * eny + wr_tot[2] * enz;
      const double wr_n[3] = {enx * wr_dot_delta, eny * wr_dot_delta, enz * wr_dot_delta};
      for (int k = 0; k < 3; k++) wr_tot[k] = wr_tot[k] - wr_n[k];
    }
    const double kr = kt * reff * reff;
    double rt[3];
    for (int k = 0; k < 3; k++) rt[k] = rhist[k] + wr_tot[k] * (dt * kr);
    const double rt_mag = sqrt(rt[0] * rt[0] + rt[1] * rt[1] + rt[2] * rt[2]);
    const double rt_max = fabs(Fn) * reff * rmu;
    if (rt_mag > rt_max) {
      const double factor = rt_max / rt_mag;
      rt[0] *= factor;
      rt[1] *= factor;
      rt[2] *= factor;
    }
    if (shearupdate)
      for (int k = 0; k < 3; k++) rhist[k] = rt[k];
    dTi[0] -= rt[0];
    dTi[1] -= rt[1];
    dTi[2] -= rt[2];
    dTj[0] += rt[0];
    dTj[1] += rt[1];
    dTj[2] += rt[2];
  } else if (mdl->rolling_cdt) {
    /* rolling model: constant directional torque */
    const double rmu = mat->mur[it][jt];
    double rt[3] = {0.0, 0.0, 0.0};
    if (c->is_wall) {
      const double wrmag = sqrt(wr1 * wr1 + wr2 * wr2 + wr3 * wr3);
      if (wrmag > 0.) {
        const double Fnw = deltan * kn;
        rt[0] = rmu * Fnw * wr1 / wrmag * reff;
        rt[1] = rmu * Fnw * wr2 / wrmag * reff;
        rt[2] = rmu * Fnw * wr3 / wrmag * reff;
        if (!mdl->torsion_torque) {
          const double d = rt[0] * enx + rt[1] * eny + rt[2] * enz;
          const double n0 = enx * d, n1 = eny * d, n2 = enz * d;
          rt[0] = rt[0] - n0;
          rt[1] = rt[1] - n1;
          rt[2] = rt[2] - n2;
        }
      }
    } else {
      const double wrr[3] = {c->wi[0] - c->wj[0], c->wi[1] - c->wj[1], c->wi[2] - c->wj[2]};
      const double mag = sqrt(wrr[0] * wrr[0] + wrr[1] * wrr[1] + wrr[2] * wrr[2]);
      if (mag > 0.) {
        const double fac = rmu * kn * deltan * reff / mag;
        rt[0] = wrr[0] * fac;
        rt[1] = wrr[1] * fac;
        rt[2] = wrr[2] * fac;
        if (!mdl->torsion_torque) {
          const double d = rt[0] * enx + rt[1] * eny + rt[2] * enz;
          const double n0 = enx * d, n1 = eny * d, n2 = enz * d;
          rt[0] = rt[0] - n0;
          rt[1] = rt[1] - n1;
          rt[2] = rt[2] - n2;
        }
      }
    }
    dTi[0] -= rt[0];
    dTi[1] -= rt[1];
    dTi[2] -= rt[2];
    dTj[0] += rt[0];
    dTj[1] += rt[1];
    dTj[2] += rt[2];
  }
}

/* Granular<CM>::compute_force inner body, pair_gran_base.h:286-430 */
int orc_contact_pair(const orc_model *mdl, const orc_materials *mat, int itype, int jtype,
                     const double *del, double radi, double radj, const double *vi,
                     const double *vj, const double *wi, const double *wj, double mi,
                     double mj, double dt, int shearupdate, double *shear, double *dFi,
                     double *dTi, double *dTj) {
  double rh[3] = {0.0, 0.0, 0.0};
  return orc_contact_pair_r(mdl, mat, itype, jtype, del, radi, radj, vi, vj, wi, wj, mi, mj, dt,
                            shearupdate, shear, rh, dFi, dTi, dTj);
}

int orc_contact_pair_r(const orc_model *mdl, const orc_materials *mat, int itype, int jtype,
                       const double *del, double radi, double radj, const double *vi,
                       const double *vj, const double *wi, const double *wj, double mi,
                       double mj, double dt, int shearupdate, double *shear, double *rhist,
                       double *dFi, double *dTi, double *dTj) {
  const double rsq = del[0] * del[0] + del[1] * del[1] + del[2] * del[2];
  const double radsum = radi + radj;
  if (!(rsq < radsum * radsum)) return 0;
  contact_in c;
  c.is_wall = 0;
  c.radi = radi;
  c.radj = radj;
  c.radsum = radsum;
  c.r = sqrt(rsq);
  c.rinv = 1.0 / c.r;
  c.en[0] = del[0] * c.rinv;
  c.en[1] = del[1] * c.rinv;
  c.en[2] = del[2] * c.rinv;
  c.delta[0] = del[0];
  c.delta[1] = del[1];
  c.delta[2] = del[2];
  c.meff = mi * mj / (mi + mj);
  c.vi = vi;
  c.vj = vj;
  c.wi = wi;
  c.wj = wj;
  c.deltan_in = 0.0;
  c.itype = itype;
  c.jtype = jtype;
  model_chain(mdl, mat, &c, dt, shearupdate, shear, rhist, dFi, dTi, dTj);
  return 1;
}

/* Walls::Granular<CM>::compute_force, fix_wall_gran_base.h:148-361 */
void orc_contact_wall(const orc_model *mdl, const orc_materials *mat, int itype, int wtype,
                      double sidata_deltan, const double *sidata_delta, double radi,
                      const double *vi, const double *wi, const double *vwall, double mi,
                      double dt, int shearupdate, double *shear, double *dFi, double *dTi) {
  double rh[3] = {0.0, 0.0, 0.0};
  orc_contact_wall_r(mdl, mat, itype, wtype, sidata_deltan, sidata_delta, radi, vi, wi, vwall, mi,
                     dt, shearupdate, shear, rh, dFi, dTi);
}

void orc_contact_wall_r(const orc_model *mdl, const orc_materials *mat, int itype, int wtype,
                        double sidata_deltan, const double *sidata_delta, double radi,
                        const double *vi, const double *wi, const double *vwall, double mi,
                        double dt, int shearupdate, double *shear, double *rhist, double *dFi,
                        double *dTi) {
  contact_in c;
  double dTj[3];
  const double zero[3] = {0.0, 0.0, 0.0};
  c.is_wall = 1;
  c.radi = radi;
  c.radj = 0.0;
  c.r = radi - sidata_deltan;
  c.rinv = 1.0 / c.r;
  c.en[0] = sidata_delta[0] * c.rinv;
  c.en[1] = sidata_delta[1] * c.rinv;
  c.en[2] = sidata_delta[2] * c.rinv;
  c.delta[0] = sidata_delta[0];
  c.delta[1] = sidata_delta[1];
  c.delta[2] = sidata_delta[2];
  c.radsum = radi;
  c.meff = mi;
  c.vi = vi;
  c.vj = vwall;
  c.wi = wi;
  c.wj = zero;
  c.deltan_in = sidata_deltan;
  c.itype = itype;
  c.jtype = wtype;
  model_chain(mdl, mat, &c, dt, shearupdate, shear, rhist, dFi, dTi, dTj);
}

/* ------------------------------------------------------------------------
 * geometry helpers
 * ---------------------------------------------------------------------- */
/* x_i - x_j with the minimum-image shift applied to the difference */
static inline void min_image_del(const orc_params *p, const double *xi, const double *xj,
                                 double *del) {
  for (int d = 0; d < 3; d++) {
    double dd = xi[d] - xj[d];
    if (p->periodic[d]) {
      const double L = p->boxhi[d] - p->boxlo[d];
      if (dd > 0.5 * L)
        dd -= L;
      else if (dd < -0.5 * L)
        dd += L;
    }
    del[d] = dd;
  }
}

/* neighbor predicate, neigh_gran.cpp:571-578 */
static inline int within_cut(const orc_params *p, const double *del, double radi,
                             double radj, double *rsq_out) {
  const double rsq = del[0] * del[0] + del[1] * del[1] + del[2] * del[2];
  const double radsum = (radi + radj) * p->cdf;
  const double cutsq = (radsum + p->skin) * (radsum + p->skin);
  *rsq_out = rsq;
  return rsq <= cutsq;
}

int64_t orc_pairs_bruteforce(int n, const double *x, const double *radius,
                             const orc_params *p, int64_t cap, int *pairs) {
  int64_t np = 0;
  for (int i = 0; i < n; i++)
    for (int j = i + 1; j < n; j++) {
      double del[3], rsq;
      min_image_del(p, &x[3 * i], &x[3 * j], del);
      if (within_cut(p, del, radius[i], radius[j], &rsq)) {
        if (np < cap) {
          pairs[2 * np] = i;
          pairs[2 * np + 1] = j;
        }
        np++;
      }
    }
  return np <= cap ? np : -np;
}

/* ------------------------------------------------------------------------
 * triangle mesh data (fix mesh/surface): per-element geometry, topology,
 * node velocities, and per-particle contact history slots
 * ---------------------------------------------------------------------- */
#define MESH_MAXNEIGH 5   /* SurfaceMesh<3,5> (tri_mesh.h) */
#define MESH_MAXSLOT 16   /* history slots per particle */
/* values per mesh history slot: the shear, then the epsd2 rolling torque */
#define MESH_HW 6
#define MESH_MAXCAND 128  /* candidate triangles per particle */
#define SMALL_TRIMESH (1.e-10)  /* tri_mesh_I.h:47-50 */
#define LARGE_TRIMESH 1000000
#define EPSILON_PRECISION 1e-8  /* multi_node_mesh.h:59 */
#define EPSILON_CURVATURE 0.00001 /* surface_mesh.h:61 */
#define SMALL_DMBRDR 1.0e-8       /* domain.h:64 */

typedef struct {
  int ntri;
  double *node, *center, *rbound, *edgevec, *edgelen, *snorm, *edgenorm, *vnode, *node_last;
  int *obtuse;
  unsigned char *eact, *cact;
  int *nneigh, *neigh;
  int wall_type;
  orc_model model;
  int moving; /* 1 rotate, 2 linear, 3 wiggle (its fix move/mesh) */
  double origin[3], axis[3], omega; /* rotate */
  double vel[3];                     /* linear */
  double ampl[3], womega;            /* wiggle */
  int vord;                          /* viblin (axis) / vibrot (origin, axis) */
  double vampl[30], vphi[30], vomega[30];
  double time;                       /* FixMoveMesh::time_ */
  int stress;                        /* mesh/surface/stress */
  double p_ref[3], ftot[3], ttot[3]; /* reference point, total force and torque on the mesh */
  double curvature, precision;
  /* per particle */
  int *ncand, *cand;
  int *npartner, *ptri;
  unsigned char *keep;
  double *phist;
} orc_mesh;

/* ------------------------------------------------------------------------
 * system
 * ---------------------------------------------------------------------- */
struct orc_sys {
  orc_params p;
  orc_materials m;
  int n;
  double *x, *v, *w, *f, *t, *r, *rm;
  int *type, *tag;
  double *xhold;
  /* half neighbor list, CSR over i (neigh_list.h pages) */
  int64_t *first;
  int *nbr, *flag;
  double *hist;
  double *rhist; /* rolling-torque history of epsd2, beside hist */
  int64_t cap;
  int built;
  /* per-atom partner arrays, fix_contact_history.cpp:305-425 */
  int *npartner;
  int64_t *poff;
  int *partner;
  double *phist;
  double *prhist;
  int64_t pcap;
  /* primitive wall history, one property/atom vector per wall (fix_wall_gran.cpp:498-522) */
  double *whist;
  double *wrhist;
  int64_t ntimestep;
  int64_t nrebuilds;
  double cutneighmax;
  int nmesh;
  orc_mesh mesh[ORC_MAXMESH];
  int64_t mesh_next_reneighbor;
  int64_t last_build; /* timestep of the last neighbor build (Neighbor::lastcall) */
  int64_t mesh_forced;
  int64_t nlost;            /* atoms beyond an f face at the last build */
  int64_t first_lost_step;  /* first timestep whose build lost atoms, -1 if none */
  double box_lo[3], box_hi[3]; /* box after the last build's reset_box */
  /* particles waiting to be inserted at the next step (orc_insert) */
  int npend;
  double *pend; /* 11 values each: x[3] v[3] omega[3] radius rmass */
  int *pend_it; /* type, tag */
  double rmax_future; /* orc_set_max_radius */
};

static void *xcalloc(size_t n, size_t sz) {
  void *p = calloc(n ? n : 1, sz);
  if (!p) abort();
  return p;
}

orc_sys *orc_create(const orc_params *p, const orc_materials *m, int n, const double *x,
                    const double *v, const double *omega, const double *radius,
                    const double *rmass, const int *type, const int *tag) {
  orc_sys *s = (orc_sys *)xcalloc(1, sizeof(orc_sys));
  s->p = *p;
  s->m = *m;
  s->n = n;
  s->first_lost_step = -1;
  for (int d = 0; d < 3; d++) {
    s->box_lo[d] = p->boxlo[d];
    s->box_hi[d] = p->boxhi[d];
  }
  s->x = (double *)xcalloc(3 * (size_t)n, sizeof(double));
  s->v = (double *)xcalloc(3 * (size_t)n, sizeof(double));
  s->w = (double *)xcalloc(3 * (size_t)n, sizeof(double));
  s->f = (double *)xcalloc(3 * (size_t)n, sizeof(double));
  s->t = (double *)xcalloc(3 * (size_t)n, sizeof(double));
  s->xhold = (double *)xcalloc(3 * (size_t)n, sizeof(double));
  s->r = (double *)xcalloc((size_t)n, sizeof(double));
  s->rm = (double *)xcalloc((size_t)n, sizeof(double));
  s->type = (int *)xcalloc((size_t)n, sizeof(int));
  s->tag = (int *)xcalloc((size_t)n, sizeof(int));
  memcpy(s->x, x, 3 * sizeof(double) * (size_t)n);
  memcpy(s->v, v, 3 * sizeof(double) * (size_t)n);
  memcpy(s->w, omega, 3 * sizeof(double) * (size_t)n);
  memcpy(s->r, radius, sizeof(double) * (size_t)n);
  memcpy(s->rm, rmass, sizeof(double) * (size_t)n);
  memcpy(s->type, type, sizeof(int) * (size_t)n);
  memcpy(s->tag, tag, sizeof(int) * (size_t)n);
  s->first = (int64_t *)xcalloc((size_t)n + 1, sizeof(int64_t));
  s->npartner = (int *)xcalloc((size_t)n, sizeof(int));
  s->poff = (int64_t *)xcalloc((size_t)n + 1, sizeof(int64_t));
  s->whist = (double *)xcalloc((size_t)n * (p->nplanes > 0 ? p->nplanes : 1) * 3, sizeof(double));
  s->wrhist = (double *)xcalloc((size_t)n * (p->nplanes > 0 ? p->nplanes : 1) * 3, sizeof(double));
  /* cutneighmax = max_ij (rmax_i + rmax_j) * cdf + skin, pair_gran.cpp:623-635, neighbor.cpp:310-332 */
  double rmax = 0.0;
  for (int i = 0; i < n; i++)
    if (radius[i] > rmax) rmax = radius[i];
  s->cutneighmax = (rmax + rmax) * p->cdf + p->skin;
  s->mesh_next_reneighbor = -1;
  return s;
}

/* PairGran::init_one takes the insertion fixes' max_rad as dynamic radii
 * (pair_gran.cpp:538-551): cutneighmax covers particles still to come */
void orc_set_max_radius(orc_sys *s, double rmax_future) {
  double rmax = rmax_future;
  for (int i = 0; i < s->n; i++)
    if (s->r[i] > rmax) rmax = s->r[i];
  s->rmax_future = rmax_future;
  s->cutneighmax = (rmax + rmax) * s->p.cdf + s->p.skin;
}

void orc_insert(orc_sys *s, int n, const double *x, const double *v, const double *omega,
                const double *radius, const double *rmass, const int *type, const int *tag) {
  s->pend = (double *)realloc(s->pend, 11 * sizeof(double) * (size_t)(s->npend + n + 1));
  s->pend_it = (int *)realloc(s->pend_it, 2 * sizeof(int) * (size_t)(s->npend + n + 1));
  if (!s->pend || !s->pend_it) abort();
  for (int k = 0; k < n; k++) {
    double *q = s->pend + 11 * (size_t)(s->npend + k);
    for (int d = 0; d < 3; d++) {
      q[d] = x[3 * k + d];
      q[3 + d] = v[3 * k + d];
      q[6 + d] = omega[3 * k + d];
    }
    q[9] = radius[k];
    q[10] = rmass[k];
    s->pend_it[2 * (s->npend + k)] = type[k];
    s->pend_it[2 * (s->npend + k) + 1] = tag[k];
  }
  s->npend += n;
}

static void *grow_zero(void *p, size_t old_bytes, size_t new_bytes) {
  unsigned char *q = (unsigned char *)realloc(p, new_bytes ? new_bytes : 1);
  if (!q) abort();
  if (new_bytes > old_bytes) memset(q + old_bytes, 0, new_bytes - old_bytes);
  return q;
}

/* ParticleToInsert::insert (particleToInsert.cpp:99-147): the queued atoms are
 * created behind the present ones with no contact history */
static void insert_pending(orc_sys *s) {
  const size_t n0 = (size_t)s->n, n1 = n0 + (size_t)s->npend, D = sizeof(double);
  const size_t np = (size_t)(s->p.nplanes > 0 ? s->p.nplanes : 1);
  s->x = (double *)grow_zero(s->x, 3 * D * n0, 3 * D * n1);
  s->v = (double *)grow_zero(s->v, 3 * D * n0, 3 * D * n1);
  s->w = (double *)grow_zero(s->w, 3 * D * n0, 3 * D * n1);
  s->f = (double *)grow_zero(s->f, 3 * D * n0, 3 * D * n1);
  s->t = (double *)grow_zero(s->t, 3 * D * n0, 3 * D * n1);
  s->xhold = (double *)grow_zero(s->xhold, 3 * D * n0, 3 * D * n1);
  s->r = (double *)grow_zero(s->r, D * n0, D * n1);
  s->rm = (double *)grow_zero(s->rm, D * n0, D * n1);
  s->type = (int *)grow_zero(s->type, sizeof(int) * n0, sizeof(int) * n1);
  s->tag = (int *)grow_zero(s->tag, sizeof(int) * n0, sizeof(int) * n1);
  s->first = (int64_t *)grow_zero(s->first, sizeof(int64_t) * (n0 + 1), sizeof(int64_t) * (n1 + 1));
  for (size_t i = n0 + 1; i <= n1; i++) s->first[i] = s->first[n0]; /* empty rows */
  s->npartner = (int *)grow_zero(s->npartner, sizeof(int) * n0, sizeof(int) * n1);
  s->poff = (int64_t *)grow_zero(s->poff, sizeof(int64_t) * (n0 + 1), sizeof(int64_t) * (n1 + 1));
  s->whist = (double *)grow_zero(s->whist, 3 * D * np * n0, 3 * D * np * n1);
  s->wrhist = (double *)grow_zero(s->wrhist, 3 * D * np * n0, 3 * D * np * n1);
  for (int k = 0; k < s->nmesh; k++) {
    orc_mesh *m = &s->mesh[k];
    m->ncand = (int *)grow_zero(m->ncand, sizeof(int) * n0, sizeof(int) * n1);
    m->cand = (int *)grow_zero(m->cand, sizeof(int) * MESH_MAXCAND * n0, sizeof(int) * MESH_MAXCAND * n1);
    m->npartner = (int *)grow_zero(m->npartner, sizeof(int) * n0, sizeof(int) * n1);
    m->ptri = (int *)grow_zero(m->ptri, sizeof(int) * MESH_MAXSLOT * n0, sizeof(int) * MESH_MAXSLOT * n1);
    for (size_t q = MESH_MAXSLOT * n0; q < MESH_MAXSLOT * n1; q++) m->ptri[q] = -1;
    m->keep = (unsigned char *)grow_zero(m->keep, MESH_MAXSLOT * n0, MESH_MAXSLOT * n1);
    m->phist = (double *)grow_zero(m->phist, MESH_HW * D * MESH_MAXSLOT * n0, MESH_HW * D * MESH_MAXSLOT * n1);
  }
  double rmax = s->rmax_future;
  for (size_t k = 0; k < (size_t)s->npend; k++) {
    const double *q = s->pend + 11 * k;
    const size_t i = n0 + k;
    for (int d = 0; d < 3; d++) {
      s->x[3 * i + d] = q[d];
      s->v[3 * i + d] = q[3 + d];
      s->w[3 * i + d] = q[6 + d];
    }
    s->r[i] = q[9];
    s->rm[i] = q[10];
    s->type[i] = s->pend_it[2 * k];
    s->tag[i] = s->pend_it[2 * k + 1];
  }
  s->n = (int)n1;
  s->npend = 0;
  for (size_t i = 0; i < n1; i++)
    if (s->r[i] > rmax) rmax = s->r[i];
  s->cutneighmax = (rmax + rmax) * s->p.cdf + s->p.skin;
}

static void mesh_free(orc_mesh *m);
static void mesh_rebuild(orc_sys *s);
static void mesh_move(orc_sys *s);
static int mesh_decide_rebuild(const orc_sys *s);
static void mesh_wall_compute(orc_sys *s, int im, int shearupdate);

void orc_destroy(orc_sys *s) {
  if (!s) return;
  for (int k = 0; k < s->nmesh; k++) mesh_free(&s->mesh[k]);
  free(s->x);
  free(s->v);
  free(s->w);
  free(s->f);
  free(s->t);
  free(s->xhold);
  free(s->r);
  free(s->rm);
  free(s->type);
  free(s->tag);
  free(s->first);
  free(s->nbr);
  free(s->flag);
  free(s->hist);
  free(s->rhist);
  free(s->npartner);
  free(s->poff);
  free(s->partner);
  free(s->phist);
  free(s->prhist);
  free(s->whist);
  free(s->wrhist);
  free(s->pend);
  free(s->pend_it);
  free(s);
}

/* Domain::pbc, domain.cpp:530-620 (orthogonal box) */
static void pbc(orc_sys *s) {
  for (int i = 0; i < s->n; i++)
    for (int d = 0; d < 3; d++) {
      if (!s->p.periodic[d]) continue;
      const double lo = s->p.boxlo[d], hi = s->p.boxhi[d], L = hi - lo;
      double *xi = &s->x[3 * i + d];
      if (*xi < lo) *xi += L;
      if (*xi >= hi) {
        *xi -= L;
        if (*xi < lo) *xi = lo;
      }
    }
}

/* Domain::reset_box (domain.cpp:381-449) for the shrink-wrapped faces: the
 * atom extent (start values BIG = 1e20, domain_definitions.h:60) plus
 * small = SMALL * created box length (SMALL 1e-4, domain.cpp:85, 254-259);
 * an m face takes the outer of that and its created position */
static void reset_box(orc_sys *s) {
  for (int d = 0; d < 3; d++) {
    s->box_lo[d] = s->p.boxlo[d];
    s->box_hi[d] = s->p.boxhi[d];
    if (s->p.periodic[d]) continue;
    const int slo = (s->p.shrink_faces >> (2 * d)) & 1, shi = (s->p.shrink_faces >> (2 * d + 1)) & 1;
    if (!slo && !shi) continue;
    double mn = 1.0e20, mx = -1.0e20;
    for (int i = 0; i < s->n; i++) {
      if (s->x[3 * i + d] < mn) mn = s->x[3 * i + d];
      if (s->x[3 * i + d] > mx) mx = s->x[3 * i + d];
    }
    const double small = 1.0e-4 * (s->p.boxhi[d] - s->p.boxlo[d]);
    if (slo) {
      s->box_lo[d] = mn - small;
      if (((s->p.shrink_min >> (2 * d)) & 1) && s->p.boxlo[d] < s->box_lo[d]) s->box_lo[d] = s->p.boxlo[d];
    }
    if (shi) {
      s->box_hi[d] = mx + small;
      if (((s->p.shrink_min >> (2 * d + 1)) & 1) && s->p.boxhi[d] > s->box_hi[d]) s->box_hi[d] = s->p.boxhi[d];
    }
  }
}

/* Comm::exchange with one process (comm.cpp:945-1003): an atom with
 * x < lo or x >= hi in a dimension is sent to itself and kept only inside the
 * box, so an atom beyond an f face is lost. Counted here (the atoms stay in the
 * arrays; Thermo::lost_check turns a loss into an error by default). */
static void count_lost(orc_sys *s) {
  int64_t lost = 0;
  for (int i = 0; i < s->n; i++) {
    int out = 0;
    for (int d = 0; d < 3; d++) {
      if (s->p.periodic[d]) continue;
      const double xd = s->x[3 * i + d];
      if (!((s->p.shrink_faces >> (2 * d)) & 1) && xd < s->box_lo[d]) out = 1;
      if (!((s->p.shrink_faces >> (2 * d + 1)) & 1) && xd >= s->box_hi[d]) out = 1;
    }
    lost += out;
  }
  s->nlost = lost;
  if (lost && s->first_lost_step < 0) s->first_lost_step = s->ntimestep;
}

int64_t orc_lost_atoms(const orc_sys *s, int64_t *first_step) {
  if (first_step) *first_step = s->first_lost_step;
  return s->nlost;
}

void orc_get_box(const orc_sys *s, double *boxlo, double *boxhi) {
  for (int d = 0; d < 3; d++) {
    boxlo[d] = s->box_lo[d];
    boxhi[d] = s->box_hi[d];
  }
}

/* FixContactHistory::pre_exchange, fix_contact_history.cpp:305-425: copy the
 * history of every flagged list entry to both partners, j side negated
 * (newtonflag "1" for shearx/y/z, tangential_model_history.h:78-80) */
static void carry_out(orc_sys *s) {
  const int n = s->n;
  memset(s->npartner, 0, sizeof(int) * (size_t)n);
  if (!s->built) {
    memset(s->poff, 0, sizeof(int64_t) * ((size_t)n + 1));
    return;
  }
  for (int i = 0; i < n; i++)
    for (int64_t e = s->first[i]; e < s->first[i + 1]; e++)
      if (s->flag[e]) {
        s->npartner[i]++;
        s->npartner[s->nbr[e]]++;
      }
  s->poff[0] = 0;
  for (int i = 0; i < n; i++) s->poff[i + 1] = s->poff[i] + s->npartner[i];
  if (s->poff[n] > s->pcap) {
    s->pcap = s->poff[n] + 1024;
    free(s->partner);
    free(s->phist);
    free(s->prhist);
    s->partner = (int *)xcalloc((size_t)s->pcap, sizeof(int));
    s->phist = (double *)xcalloc(3 * (size_t)s->pcap, sizeof(double));
    s->prhist = (double *)xcalloc(3 * (size_t)s->pcap, sizeof(double));
  }
  memset(s->npartner, 0, sizeof(int) * (size_t)n);
  for (int i = 0; i < n; i++)
    for (int64_t e = s->first[i]; e < s->first[i + 1]; e++)
      if (s->flag[e]) {
        const int j = s->nbr[e];
        int64_t m = s->poff[i] + s->npartner[i]++;
        s->partner[m] = s->tag[j];
        for (int d = 0; d < 3; d++) s->phist[3 * m + d] = s->hist[3 * e + d];
        for (int d = 0; d < 3; d++) s->prhist[3 * m + d] = s->rhist[3 * e + d];
        m = s->poff[j] + s->npartner[j]++;
        s->partner[m] = s->tag[i];
        /* every history value is registered with newtonflag 1: the partner's
         * copy is negated (fix_contact_history.cpp:400-410) */
        for (int d = 0; d < 3; d++) s->phist[3 * m + d] = -s->hist[3 * e + d];
        for (int d = 0; d < 3; d++) s->prhist[3 * m + d] = -s->rhist[3 * e + d];
      }
}

/* cell grid for the binned build: cells no smaller than cutneighmax so that
 * the 27-cell stencil covers the cutoff (the reference uses half-size bins and
 * a matching stencil, neighbor.cpp:1642-1810, neigh_stencil.cpp:116-130; the
 * stored pair set is the same) */
typedef struct {
  int nc[3];
  double lo[3], inv[3];
} grid_t;

static void make_grid(const orc_sys *s, grid_t *g) {
  for (int d = 0; d < 3; d++) {
    const double L = s->p.boxhi[d] - s->p.boxlo[d];
    int nc = (int)(L / s->cutneighmax);
    if (nc < 1) nc = 1;
    if (s->p.periodic[d] && nc < 3) nc = 1;
    if (nc > 1024) nc = 1024;
    g->nc[d] = nc;
    g->lo[d] = s->p.boxlo[d];
    g->inv[d] = nc / L;
  }
}

static int cell_coord(const grid_t *g, int d, double x) {
  int c = (int)floor((x - g->lo[d]) * g->inv[d]);
  if (c < 0) c = 0;
  if (c >= g->nc[d]) c = g->nc[d] - 1;
  return c;
}

/* Neighbor::build + granular_bin_no_newton, neighbor.cpp:1469-1560 and
 * neigh_gran.cpp:485-639, including history carry-in by partner tag scan */
static void build(orc_sys *s) {
  const int n = s->n;
  grid_t g;
  make_grid(s, &g);
  const int ncell = g.nc[0] * g.nc[1] * g.nc[2];
  int *head = (int *)malloc(sizeof(int) * (size_t)ncell);
  int *next = (int *)malloc(sizeof(int) * (size_t)(n ? n : 1));
  int *cellof = (int *)malloc(sizeof(int) * (size_t)(n ? n : 1));
  for (int c = 0; c < ncell; c++) head[c] = -1;
  for (int i = n - 1; i >= 0; i--) {
    const int cx = cell_coord(&g, 0, s->x[3 * i]);
    const int cy = cell_coord(&g, 1, s->x[3 * i + 1]);
    const int cz = cell_coord(&g, 2, s->x[3 * i + 2]);
    const int c = cx + g.nc[0] * (cy + g.nc[1] * cz);
    cellof[i] = c;
    next[i] = head[c];
    head[c] = i;
  }
  /* xhold, neighbor.cpp:1480-1490 */
  memcpy(s->xhold, s->x, 3 * sizeof(double) * (size_t)n);

  int64_t np = 0;
  for (int pass = 0; pass < 2; pass++) {
    np = 0;
    for (int i = 0; i < n; i++) {
      if (pass == 1) s->first[i] = np;
      const int c = cellof[i];
      const int ci[3] = {c % g.nc[0], (c / g.nc[0]) % g.nc[1], c / (g.nc[0] * g.nc[1])};
      int lo[3], hi[3];
      for (int d = 0; d < 3; d++) {
        if (g.nc[d] == 1) {
          lo[d] = hi[d] = 0;
        } else if (s->p.periodic[d]) {
          lo[d] = -1;
          hi[d] = 1;
        } else {
          lo[d] = ci[d] > 0 ? -1 : 0;
          hi[d] = ci[d] < g.nc[d] - 1 ? 1 : 0;
        }
      }
      for (int dz = lo[2]; dz <= hi[2]; dz++)
        for (int dy = lo[1]; dy <= hi[1]; dy++)
          for (int dx = lo[0]; dx <= hi[0]; dx++) {
            const int cx = (ci[0] + dx + g.nc[0]) % g.nc[0];
            const int cy = (ci[1] + dy + g.nc[1]) % g.nc[1];
            const int cz = (ci[2] + dz + g.nc[2]) % g.nc[2];
            for (int j = head[cx + g.nc[0] * (cy + g.nc[1] * cz)]; j >= 0; j = next[j]) {
              if (j <= i) continue;
              double del[3], rsq;
              min_image_del(&s->p, &s->x[3 * i], &s->x[3 * j], del);
              if (!within_cut(&s->p, del, s->r[i], s->r[j], &rsq)) continue;
              if (pass == 1) {
                s->nbr[np] = j;
                const double radsum = (s->r[i] + s->r[j]) * s->p.cdf;
                double *h = &s->hist[3 * np];
                double *rh = &s->rhist[3 * np];
                int fl = 0;
                h[0] = h[1] = h[2] = 0.0;
                rh[0] = rh[1] = rh[2] = 0.0;
                if (rsq < radsum * radsum) {
                  for (int64_t m = s->poff[i]; m < s->poff[i] + s->npartner[i]; m++)
                    if (s->partner[m] == s->tag[j]) {
                      fl = 1;
                      h[0] = s->phist[3 * m];
                      h[1] = s->phist[3 * m + 1];
                      h[2] = s->phist[3 * m + 2];
                      rh[0] = s->prhist[3 * m];
                      rh[1] = s->prhist[3 * m + 1];
                      rh[2] = s->prhist[3 * m + 2];
                      break;
                    }
                }
                s->flag[np] = fl;
              }
              np++;
            }
          }
    }
    if (pass == 0 && np > s->cap) {
      s->cap = np + np / 4 + 1024;
      free(s->nbr);
      free(s->flag);
      free(s->hist);
      free(s->rhist);
      s->nbr = (int *)xcalloc((size_t)s->cap, sizeof(int));
      s->flag = (int *)xcalloc((size_t)s->cap, sizeof(int));
      s->hist = (double *)xcalloc(3 * (size_t)s->cap, sizeof(double));
      s->rhist = (double *)xcalloc(3 * (size_t)s->cap, sizeof(double));
    }
  }
  s->first[n] = np;
  s->built = 1;
  s->nrebuilds++;
  free(head);
  free(next);
  free(cellof);
}

/* full rebuild sequence of Verlet::run, verlet.cpp:313-336 */
static void rebuild(orc_sys *s) {
  s->last_build = s->ntimestep; /* Neighbor::build resets ago (neighbor.cpp) */
  carry_out(s);
  pbc(s);
  reset_box(s);  /* box_change with shrink-wrapped faces (verlet.cpp:322-325) */
  count_lost(s);
  build(s);
  mesh_rebuild(s);
}

/* Neighbor::check_distance, neighbor.cpp:1385-1461 (no box change, radius fixed) */
static int check_distance(const orc_sys *s) {
  const double deltasq = 0.25 * s->p.skin * s->p.skin;
  int flag = 0;
  for (int i = 0; i < s->n; i++) {
    const double dx = s->x[3 * i] - s->xhold[3 * i];
    const double dy = s->x[3 * i + 1] - s->xhold[3 * i + 1];
    const double dz = s->x[3 * i + 2] - s->xhold[3 * i + 2];
    const double rsq = dx * dx + dy * dy + dz * dz;
    if (rsq > deltasq) flag = 1;
  }
  return flag;
}

/* Neighbor::decide, neighbor.cpp:1357-1372, with delay 0: a build every
 * `every` steps since the last one, if check is off or an atom moved more
 * than skin/2 */
static int decide(const orc_sys *s) {
  const int every = s->p.neigh_every > 0 ? s->p.neigh_every : 1;
  const int64_t ago = s->ntimestep - s->last_build;
  if (ago % every != 0) return 0;
  if (s->p.neigh_no_check) return 1;
  return check_distance(s);
}

/* Bits of a list entry's contact flag (contact_model_constants.h:64-68). A
 * build stores 1 (the normal bit) for a carried pair (neigh_gran.cpp:592). */
#define FLAG_NORMAL 1
#define FLAG_COHESION 2
#define FLAG_TANGENTIAL 4
#define FLAG_ROLLING 8

/* the normal models whose surfacesIntersect / surfacesClose set the normal
 * bit (normal_model_hertz.h:204,409; normal_model_hooke.h:229,440;
 * normal_model_hooke_stiffness.h:196,369); hertz/stiffness sets none */
static int sets_normal_bit(const orc_model *m) {
  return m->normal_model != ORC_NORMAL_HERTZ_STIFFNESS;
}

/* ContactModel::surfacesIntersect's flag bits (tangential_model_history.h:137,
 * tangential_model_no_history.h:113, cohesion_model_sjkr.h:102,
 * rolling_model_epsd2.h:103) */
static int touch_flag_bits(const orc_model *m) {
  int b = FLAG_TANGENTIAL;
  if (sets_normal_bit(m)) b |= FLAG_NORMAL;
  if (m->cohesion_model) b |= FLAG_COHESION;
  if (m->rolling_epsd2) b |= FLAG_ROLLING;
  return b;
}

/* ContactModel::surfacesClose (contact_models.h:246-253) for a pair within
 * contact_distance_factor * radsum that does not touch: the normal models
 * above set their bit, sjkr, both tangential models and epsd2 clear theirs
 * (cohesion_model_sjkr.h:129-133, tangential_model_no_history.h:244-247,
 * tangential_model_history.h:358-369, rolling_model_epsd2.h:144-151); the
 * shear history and the epsd2 rolling torque are zeroed whatever shearupdate
 * says */
static int close_flag(const orc_model *m, int flag) {
  if (sets_normal_bit(m)) flag |= FLAG_NORMAL;
  flag &= ~(FLAG_COHESION | FLAG_TANGENTIAL | FLAG_ROLLING);
  return flag;
}

/* pair loop, pair_gran_base.h:255-430 (newton off: j always local here) */
static void pair_compute(orc_sys *s, int shearupdate) {
  /* contactDistanceMultiplier, pair_gran_base.h:224 */
  const double cdm = s->p.cdf * s->p.cdf;
  for (int i = 0; i < s->n; i++)
    for (int64_t e = s->first[i]; e < s->first[i + 1]; e++) {
      const int j = s->nbr[e];
      double del[3], dFi[3], dTi[3], dTj[3];
      min_image_del(&s->p, &s->x[3 * i], &s->x[3 * j], del);
      if (orc_contact_pair_r(&s->p.model, &s->m, s->type[i], s->type[j], del, s->r[i], s->r[j],
                             &s->v[3 * i], &s->v[3 * j], &s->w[3 * i], &s->w[3 * j], s->rm[i],
                             s->rm[j], s->p.dt, shearupdate, &s->hist[3 * e], &s->rhist[3 * e],
                             dFi, dTi, dTj)) {
        s->flag[e] |= touch_flag_bits(&s->p.model);
        for (int d = 0; d < 3; d++) {
          s->f[3 * i + d] += dFi[d];
          s->t[3 * i + d] += dTi[d];
          s->f[3 * j + d] += -dFi[d];
          s->t[3 * j + d] += dTj[d];
        }
        continue;
      }
      /* surfacesClose branch, pair_gran_base.h:414-417 (reachable only with
       * contact_distance_factor > 1) */
      const double rsq = del[0] * del[0] + del[1] * del[1] + del[2] * del[2];
      const double radsum = s->r[i] + s->r[j];
      if (rsq < cdm * radsum * radsum) {
        s->flag[e] = close_flag(&s->p.model, s->flag[e]);
        if (s->p.model.tangential_model == 0)
          s->hist[3 * e] = s->hist[3 * e + 1] = s->hist[3 * e + 2] = 0.0;
        if (s->p.model.rolling_epsd2)
          s->rhist[3 * e] = s->rhist[3 * e + 1] = s->rhist[3 * e + 2] = 0.0;
      }
    }
}

/* FixWallGran::post_force_primitive, fix_wall_gran.cpp:1024-1153, with
 * Plane<dim>::resolveContact, primitive_wall_definitions.h:130-142 */
static void wall_compute(orc_sys *s, int shearupdate) {
  const double zero[3] = {0.0, 0.0, 0.0};
  for (int k = 0; k < s->p.nplanes; k++) {
    const int ax = s->p.plane_axis[k];
    const double param = s->p.plane_pos[k];
    for (int i = 0; i < s->n; i++) {
      double *h = &s->whist[3 * ((size_t)i * s->p.nplanes + k)];
      double *rh = &s->wrhist[3 * ((size_t)i * s->p.nplanes + k)];
      const double *xi = &s->x[3 * i];
      double delta[3] = {0.0, 0.0, 0.0};
      double deltan;
      if (xi[ax] > param) {
        delta[ax] = param - xi[ax];
        deltan = xi[ax] - param - s->r[i];
      } else {
        delta[ax] = param - xi[ax];
        deltan = param - xi[ax] - s->r[i];
      }
      if (deltan > s->cutneighmax) continue;
      if (deltan <= 0 || deltan < (s->p.cdf - 1.0) * s->r[i]) {
        if (!(deltan <= 0)) {
          /* compute_force with intersectflag false: surfacesClose zeroes the
           * shear (tangential_model_history.h:358-369) and the epsd2 rolling
           * torque (rolling_model_epsd2.h:144-151); no force */
          if (s->p.plane_model[k].tangential_model == 0) h[0] = h[1] = h[2] = 0.0;
          if (s->p.plane_model[k].rolling_epsd2) rh[0] = rh[1] = rh[2] = 0.0;
          continue;
        }
        const double sdelta[3] = {-delta[0], -delta[1], -delta[2]};
        double dF[3], dT[3];
        orc_contact_wall_r(&s->p.plane_model[k], &s->m, s->type[i], s->p.plane_type[k], -deltan,
                           sdelta, s->r[i], &s->v[3 * i], &s->w[3 * i], zero, s->rm[i], s->p.dt,
                           shearupdate, h, rh, dF, dT);
        for (int d = 0; d < 3; d++) {
          s->f[3 * i + d] += dF[d];
          s->t[3 * i + d] += dT[d];
        }
      } else {
        h[0] = h[1] = h[2] = 0.0;
        rh[0] = rh[1] = rh[2] = 0.0;
      }
    }
  }
}

/* FixGravity::post_force, fix_gravity.cpp:301-350 */
static void gravity(orc_sys *s) {
  if (!s->p.has_gravity) return;
  for (int i = 0; i < s->n; i++) {
    const double massone = s->rm[i];
    s->f[3 * i] += massone * s->p.grav_acc[0];
    s->f[3 * i + 1] += massone * s->p.grav_acc[1];
    s->f[3 * i + 2] += massone * s->p.grav_acc[2];
  }
}

/* post_force fixes in script order: gravity before or after the wall fixes;
 * plane walls before mesh walls */
static void forces(orc_sys *s, int shearupdate) {
  memset(s->f, 0, 3 * sizeof(double) * (size_t)s->n);
  memset(s->t, 0, 3 * sizeof(double) * (size_t)s->n);
  pair_compute(s, shearupdate);
  if (s->p.gravity_first) gravity(s);
  wall_compute(s, shearupdate);
  for (int im = 0; im < s->nmesh; im++) mesh_wall_compute(s, im, shearupdate);
  if (!s->p.gravity_first) gravity(s);
}

void orc_compute_forces(orc_sys *s, int shearupdate) { forces(s, shearupdate); }

/* FixNVESphere::initial_integrate, fix_nve_sphere.cpp:138-206 */
static void initial_integrate(orc_sys *s) {
  const double dtv = s->p.dt, dtf = 0.5 * s->p.dt * 1.0;
  const double dtfrotate = dtf / 0.4;
  for (int i = 0; i < s->n; i++) {
    const double dtfm = dtf / (s->rm[i] * 1.0);
    for (int d = 0; d < 3; d++) s->v[3 * i + d] += dtfm * s->f[3 * i + d];
    for (int d = 0; d < 3; d++) s->x[3 * i + d] += dtv * s->v[3 * i + d];
    const double dtirotate = dtfrotate / (s->r[i] * s->r[i] * s->rm[i]);
    for (int d = 0; d < 3; d++) s->w[3 * i + d] += dtirotate * s->t[3 * i + d];
  }
}

/* FixNVESphere::final_integrate, fix_nve_sphere.cpp:208-246 */
static void final_integrate(orc_sys *s) {
  const double dtf = 0.5 * s->p.dt * 1.0;
  const double dtfrotate = dtf / 0.4;
  for (int i = 0; i < s->n; i++) {
    const double dtfm = dtf / (s->rm[i] * 1.0);
    for (int d = 0; d < 3; d++) s->v[3 * i + d] += dtfm * s->f[3 * i + d];
    const double dtirotate = dtfrotate / (s->r[i] * s->r[i] * s->rm[i]);
    for (int d = 0; d < 3; d++) s->w[3 * i + d] += dtirotate * s->t[3 * i + d];
  }
}

/* Verlet::setup, verlet.cpp:134-205: carry history out, wrap, rebuild, and
 * compute forces with shearupdate = 0 (pair_gran.cpp:665-667) */
void orc_setup(orc_sys *s) {
  rebuild(s);
  /* FixMoveMesh::setup zeroes the node velocities (fix_move_mesh.cpp:211-216) */
  for (int im = 0; im < s->nmesh; im++)
    memset(s->mesh[im].vnode, 0, 9 * sizeof(double) * (size_t)s->mesh[im].ntri);
  forces(s, 0);
}

/* Verlet::run, verlet.cpp:264-392 */
void orc_run(orc_sys *s, int nsteps) {
  for (int k = 0; k < nsteps; k++) {
    s->ntimestep++;
    initial_integrate(s);
    mesh_move(s);
    /* Neighbor::decide with delay 0, every 1: forced by a fix's next_reneighbor
     * (moving mesh, or an insertion, fix_insert.cpp:311-319), else check every
     * step; the insertion is a pre_exchange fix of the build (verlet.cpp:313-320) */
    const int inserting = s->npend > 0;
    if (inserting) insert_pending(s);
    const int forced = s->mesh_next_reneighbor == s->ntimestep || inserting;
    if (forced || decide(s)) {
      if (s->mesh_next_reneighbor == s->ntimestep) s->mesh_forced++;
      rebuild(s);
    } else if (mesh_decide_rebuild(s)) {
      /* FixMesh::pre_force on a regular step (fix_mesh.cpp:572-582) */
      s->mesh_next_reneighbor = s->ntimestep + 1;
    }
    forces(s, 1);
    final_integrate(s);
  }
}

void orc_get_state(const orc_sys *s, double *x, double *v, double *omega, double *f,
                   double *torque) {
  const size_t b = 3 * sizeof(double) * (size_t)s->n;
  if (x) memcpy(x, s->x, b);
  if (v) memcpy(v, s->v, b);
  if (omega) memcpy(omega, s->w, b);
  if (f) memcpy(f, s->f, b);
  if (torque) memcpy(torque, s->t, b);
}

int64_t orc_get_pairs(const orc_sys *s, int64_t cap, int *tags, int *flags, double *hist) {
  if (!s->built) return 0;
  const int64_t np = s->first[s->n];
  if (np > cap) return -np;
  for (int i = 0; i < s->n; i++)
    for (int64_t e = s->first[i]; e < s->first[i + 1]; e++) {
      const int j = s->nbr[e];
      const int ti = s->tag[i], tj = s->tag[j];
      const double sg = ti < tj ? 1.0 : -1.0;
      tags[2 * e] = ti < tj ? ti : tj;
      tags[2 * e + 1] = ti < tj ? tj : ti;
      flags[e] = s->flag[e] != 0; /* the flag's bits collapse to set / unset */
      for (int d = 0; d < 3; d++) hist[3 * e + d] = sg * s->hist[3 * e + d];
    }
  return np;
}

/* ComputePairGranLocal::add_pair (compute_pair_gran_local.cpp:420-633) for
 * the default columns (pos, vel, id, force, torque, history), called from
 * PairGran::compute_pgl (pair_gran.cpp:676-708): the pair loop with
 * shearupdate 0 (pair_gran_base.h:275-447), which calls cpl_add_pair for every
 * pair of the half list with the force and torque the contact puts on i
 * (i_forces, zero when the pair does not touch; pair_gran.cpp:798-813). The
 * history is the pair's stored one. Rows in the order of orc_get_pairs; each
 * pair is oriented from its lower tag (the reference orients it from its
 * first local index): for a pair whose list owner has the higher tag the row
 * takes the partner's view, force -dFi and torque dTj, negated history.
 * Columns: x_lo[3], x_hi[3] (unwrapped, as stored), v_lo[3], v_hi[3], F[3],
 * T[3], periodic flag (1 if the minimum-image shift moved the pair), r_lo,
 * r_hi, shear history[3], and with epsd2 the rolling-torque history[3]. Nothing in the
 * system changes. Returns the number of pairs (negated when above cap). */
int64_t orc_pair_local(const orc_sys *s, int64_t cap, int *tags, double *data) {
  if (!s->built) return 0;
  const int64_t np = s->first[s->n];
  if (np > cap) return -np;
  const int ncol = 21 + (s->p.model.rolling_epsd2 ? 6 : 3);
  for (int i = 0; i < s->n; i++)
    for (int64_t e = s->first[i]; e < s->first[i + 1]; e++) {
      const int j = s->nbr[e];
      double del[3], dFi[3] = {0, 0, 0}, dTi[3] = {0, 0, 0}, dTj[3] = {0, 0, 0};
      double sh[3], rh[3];
      for (int d = 0; d < 3; d++) {
        sh[d] = s->hist[3 * e + d];
        rh[d] = s->rhist ? s->rhist[3 * e + d] : 0.0;
      }
      min_image_del(&s->p, &s->x[3 * i], &s->x[3 * j], del);
      orc_contact_pair_r(&s->p.model, &s->m, s->type[i], s->type[j], del, s->r[i], s->r[j], &s->v[3 * i],
                         &s->v[3 * j], &s->w[3 * i], &s->w[3 * j], s->rm[i], s->rm[j], s->p.dt, 0, sh, rh, dFi,
                         dTi, dTj);
      const int lo_first = s->tag[i] < s->tag[j];
      const int a = lo_first ? i : j, b = lo_first ? j : i;
      const double sg = lo_first ? 1.0 : -1.0;
      double *row = data + e * ncol;
      tags[2 * e] = s->tag[a];
      tags[2 * e + 1] = s->tag[b];
      int periodic = 0;
      for (int d = 0; d < 3; d++) {
        row[d] = s->x[3 * a + d];
        row[3 + d] = s->x[3 * b + d];
        row[6 + d] = s->v[3 * a + d];
        row[9 + d] = s->v[3 * b + d];
        row[12 + d] = lo_first ? dFi[d] : -dFi[d];
        row[15 + d] = lo_first ? dTi[d] : dTj[d];
        if (del[d] != s->x[3 * i + d] - s->x[3 * j + d]) periodic = 1;
      }
      row[18] = periodic;
      row[19] = s->r[a];
      row[20] = s->r[b];
      const int fl = s->flag[e] != 0;
      for (int d = 0; d < 3; d++) row[21 + d] = fl ? sg * s->hist[3 * e + d] : 0.0;
      if (s->p.model.rolling_epsd2)
        for (int d = 0; d < 3; d++) row[24 + d] = fl ? sg * s->rhist[3 * e + d] : 0.0;
    }
  return np;
}

void orc_get_wall_hist(const orc_sys *s, double *hist) {
  memcpy(hist, s->whist, sizeof(double) * 3 * (size_t)s->n * (size_t)(s->p.nplanes > 0 ? s->p.nplanes : 1));
}

int64_t orc_get_pair_rolling(const orc_sys *s, int64_t cap, double *rhist) {
  const int64_t np = s->first[s->n];
  if (np > cap) return -np;
  for (int i = 0; i < s->n; i++)
    for (int64_t e = s->first[i]; e < s->first[i + 1]; e++) {
      const double sg = s->tag[i] < s->tag[s->nbr[e]] ? 1.0 : -1.0;
      for (int d = 0; d < 3; d++) rhist[3 * e + d] = sg * s->rhist[3 * e + d];
    }
  return np;
}

void orc_get_wall_rolling(const orc_sys *s, double *rhist) {
  memcpy(rhist, s->wrhist, sizeof(double) * 3 * (size_t)s->n * (size_t)(s->p.nplanes > 0 ? s->p.nplanes : 1));
}

int64_t orc_nrebuilds(const orc_sys *s) { return s->nrebuilds; }
int64_t orc_ntimestep(const orc_sys *s) { return s->ntimestep; }
int orc_natoms(const orc_sys *s) { return s->n; }

void orc_contact_counts(const orc_sys *s, int64_t *npair_touch, int64_t *nwall_touch) {
  int64_t c = 0, w = 0;
  if (s->built)
    for (int i = 0; i < s->n; i++)
      for (int64_t e = s->first[i]; e < s->first[i + 1]; e++) {
        const int j = s->nbr[e];
        double del[3];
        min_image_del(&s->p, &s->x[3 * i], &s->x[3 * j], del);
        const double rsq = del[0] * del[0] + del[1] * del[1] + del[2] * del[2];
        const double rs = s->r[i] + s->r[j];
        if (rsq < rs * rs) c++;
      }
  for (int k = 0; k < s->p.nplanes; k++)
    for (int i = 0; i < s->n; i++) {
      const int ax = s->p.plane_axis[k];
      const double d = fabs(s->x[3 * i + ax] - s->p.plane_pos[k]) - s->r[i];
      if (d <= 0) w++;
    }
  *npair_touch = c;
  *nwall_touch = w;
}

/* =========================================================================
 * triangle-mesh walls
 *   element geometry      surface_mesh_I.h:361-447, multi_node_mesh_I.h:139-172
 *   topology              surface_mesh_I.h:454-560 (buildNeighbours),
 *                         :844-1000 (coplanarity, overlap), :1035-1190 (edges,
 *                         corners), multi_node_mesh_I.h:246-355 (node sharing)
 *   sphere contact        tri_mesh_I.h:65-251, math_extra_liggghts.h:559-570
 *   contact history       fix_contact_history_mesh_I.h:51-200,
 *                         fix_contact_history_mesh.cpp:272-507
 *   force loop            fix_wall_gran.cpp:845-1018
 *   rotation              mesh_mover_rotation.cpp:66-139, multi_node_mesh_I.h:620-660,
 *                         math_extra_liggghts.h:431-470, math_extra.h:596-602
 *   rebuild trigger       multi_node_mesh_I.h:792-825, fix_mesh.cpp:564-592
 * ========================================================================= */

static int comp_double(double a, double b, double prec) {
  if (a == b) return 1;
  if (b == 0) return a < prec && a > -prec;
  const double x = (a - b);
  return x < prec && x > -prec;
}

static inline void v3_sub(const double *a, const double *b, double *r) {
  r[0] = a[0] - b[0];
  r[1] = a[1] - b[1];
  r[2] = a[2] - b[2];
}
static inline double v3_dot(const double *a, const double *b) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}
static inline double v3_mag(const double *a) { return sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]); }
static inline void v3_cross(const double *a, const double *b, double *r) {
  r[0] = a[1] * b[2] - a[2] * b[1];
  r[1] = a[2] * b[0] - a[0] * b[2];
  r[2] = a[0] * b[1] - a[1] * b[0];
}
static inline void v3_div(double *v, double sc) {
  const double sinv = 1. / sc;
  v[0] = sinv * v[0];
  v[1] = sinv * v[1];
  v[2] = sinv * v[2];
}

static void quatquat(const double *a, const double *b, double *c) {
  c[0] = a[0] * b[0] - a[1] * b[1] - a[2] * b[2] - a[3] * b[3];
  c[1] = a[0] * b[1] + b[0] * a[1] + a[2] * b[3] - a[3] * b[2];
  c[2] = a[0] * b[2] + b[0] * a[2] + a[3] * b[1] - a[1] * b[3];
  c[3] = a[0] * b[3] + b[0] * a[3] + a[1] * b[2] - a[2] * b[1];
}

/* v <- q v q* */
static void vec_quat_rotate(double *v, const double *q) {
  const double vq[4] = {0., v[0], v[1], v[2]};
  const double qc[4] = {q[0], -q[1], -q[2], -q[3]};
  double t[4], r[4];
  quatquat(q, vq, t);
  quatquat(t, qc, r);
  v[0] = r[1];
  v[1] = r[2];
  v[2] = r[3];
}

/* edge vectors/lengths, surface normal, edge normals, obtuse index */
static void mesh_surf_props(orc_mesh *m, int t) {
  const double *n = &m->node[9 * t];
  double *ev = &m->edgevec[9 * t], *el = &m->edgelen[3 * t], *sn = &m->snorm[3 * t];
  double *en = &m->edgenorm[9 * t];
  for (int i = 0; i < 3; i++) {
    v3_sub(&n[3 * ((i + 1) % 3)], &n[3 * i], &ev[3 * i]);
    el[i] = v3_mag(&ev[3 * i]);
    v3_div(&ev[3 * i], el[i]);
  }
  v3_cross(&ev[0], &ev[3], sn);
  if (v3_mag(sn) < 1e-15) {
    v3_cross(&ev[3], &ev[6], sn);
    if (v3_mag(sn) < 1e-15) {
      v3_cross(&ev[6], &ev[0], sn);
      if (v3_mag(sn) < 1e-15) {
        const double tmpvec[] = {1.1233, 2.123231, -3.3343434};
        v3_cross(&ev[0], tmpvec, sn);
        if (v3_mag(sn) < 1e-15) {
          v3_cross(&ev[3], tmpvec, sn);
          if (v3_mag(sn) < 1e-15) {
            const double tmpvec2[] = {1.1233, -2.123231, 3.3343434};
            v3_cross(&ev[0], tmpvec2, sn);
            if (v3_mag(sn) < 1e-15) v3_cross(&ev[3], tmpvec2, sn);
          }
        }
      }
    }
  }
  v3_div(sn, v3_mag(sn));
  for (int i = 0; i < 3; i++) {
    v3_cross(&ev[3 * i], sn, &en[3 * i]);
    if (v3_mag(&en[3 * i]) < 1e-15) {
      const int o = (i + 1) % 3;
      en[3 * i] = ev[3 * o];
      en[3 * i + 1] = ev[3 * o + 1];
      en[3 * i + 2] = ev[3 * o + 2];
    } else {
      v3_div(&en[3 * i], v3_mag(&en[3 * i]));
    }
  }
  /* calcObtuseAngleIndex is called for every node and each call overwrites
   * the element's value, so the last node decides (surface_mesh_I.h:438-447) */
  for (int i = 0; i < 3; i++) {
    const double dot = v3_dot(&ev[3 * i], &ev[3 * ((i - 1 + 3) % 3)]);
    m->obtuse[t] = dot > 0. ? i : -1;
  }
}

static int nodes_equal(const orc_mesh *m, const double *a, const double *b) {
  for (int k = 0; k < 3; k++)
    if (!comp_double(a[k], b[k], m->precision)) return 0;
  return 1;
}

static int n_shared_nodes(const orc_mesh *m, int i, int j) {
  double d[3];
  v3_sub(&m->center[3 * i], &m->center[3 * j], d);
  const double radsum = m->rbound[i] + m->rbound[j];
  if (d[0] * d[0] + d[1] * d[1] + d[2] * d[2] > radsum * radsum) return 0;
  int ns = 0;
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++)
      if (nodes_equal(m, &m->node[9 * i + 3 * a], &m->node[9 * j + 3 * b])) ns++;
  return ns;
}

static int share_edge(const orc_mesh *m, int i, int j, int *iEdge, int *jEdge) {
  double d[3];
  v3_sub(&m->center[3 * i], &m->center[3 * j], d);
  const double radsum = m->rbound[i] + m->rbound[j];
  if (d[0] * d[0] + d[1] * d[1] + d[2] * d[2] > radsum * radsum) return 0;
  int n = 0, in1 = -1, jn1 = -1, in2 = -1, jn2 = -1;
  for (int a = 0; a < 3 && n < 2; a++)
    for (int b = 0; b < 3; b++)
      if (nodes_equal(m, &m->node[9 * i + 3 * a], &m->node[9 * j + 3 * b])) {
        if (n == 0) {
          in1 = a;
          jn1 = b;
          n = 1;
        } else {
          in2 = a;
          jn2 = b;
          n = 2;
          break;
        }
      }
  if (n < 2) return 0;
  *iEdge = (in1 + in2 == 2) ? 2 : (in1 < in2 ? in1 : in2);
  *jEdge = (jn1 + jn2 == 2) ? 2 : (jn1 < jn2 ? jn1 : jn2);
  return 1;
}

static int coplanar_neighs_overlap(const orc_mesh *m, int i, int iEdge, int j, int jEdge) {
  const double *pRef = &m->node[9 * i + 3 * iEdge];
  const double *edgeN = &m->edgenorm[9 * i + 3 * iEdge];
  double vI[3], vJ[3];
  v3_sub(&m->node[9 * i + 3 * ((iEdge + 2) % 3)], pRef, vI);
  v3_sub(&m->node[9 * j + 3 * ((jEdge + 2) % 3)], pRef, vJ);
  const double d1 = v3_dot(vI, edgeN), d2 = v3_dot(vJ, edgeN);
  return d1 * d2 > 0.;
}

static void handle_shared_edge(orc_mesh *m, int i, int iEdge, int j, int jEdge, int coplanar) {
  if (m->nneigh[i] < MESH_MAXNEIGH) m->neigh[MESH_MAXNEIGH * i + m->nneigh[i]] = j;
  if (m->nneigh[j] < MESH_MAXNEIGH) m->neigh[MESH_MAXNEIGH * j + m->nneigh[j]] = i;
  m->nneigh[i]++;
  m->nneigh[j]++;
  if (!coplanar || coplanar_neighs_overlap(m, i, iEdge, j, jEdge)) {
    if (i < j) {
      m->eact[3 * i + iEdge] = 0;
      m->eact[3 * j + jEdge] = 1;
    } else {
      m->eact[3 * i + iEdge] = 1;
      m->eact[3 * j + jEdge] = 0;
    }
  } else {
    m->eact[3 * i + iEdge] = 0;
    m->eact[3 * j + jEdge] = 0;
  }
}

static int contains_node(const orc_mesh *m, int t, const double *p) {
  for (int a = 0; a < 3; a++) {
    const double *q = &m->node[9 * t + 3 * a];
    if (comp_double(q[0], p[0], m->precision) && comp_double(q[1], p[1], m->precision) &&
        comp_double(q[2], p[2], m->precision))
      return a;
  }
  return -1;
}

typedef struct {
  int *visited, nvisited;
  int *has, nhas;
  double *edges, *ends; /* 2 per element with the node */
  int anyActive;
} corner_scan;

static void check_node_recursive(const orc_mesh *m, int t, const double *p, corner_scan *c) {
  for (int k = 0; k < c->nvisited; k++)
    if (c->visited[k] == t) return;
  c->visited[c->nvisited++] = t;
  const int a = contains_node(m, t, p);
  if (a < 0) return;
  c->has[c->nhas] = t;
  const int am = (a - 1 + 3) % 3;
  memcpy(&c->edges[3 * (2 * c->nhas)], &m->edgevec[9 * t + 3 * a], 3 * sizeof(double));
  memcpy(&c->edges[3 * (2 * c->nhas + 1)], &m->edgevec[9 * t + 3 * am], 3 * sizeof(double));
  memcpy(&c->ends[3 * (2 * c->nhas)], &m->node[9 * t + 3 * ((a + 1) % 3)], 3 * sizeof(double));
  memcpy(&c->ends[3 * (2 * c->nhas + 1)], &m->node[9 * t + 3 * am], 3 * sizeof(double));
  c->nhas++;
  if (m->eact[3 * t + a])
    c->anyActive = 1;
  else if (m->eact[3 * t + am])
    c->anyActive = 1;
  for (int k = 0; k < m->nneigh[t] && k < MESH_MAXNEIGH; k++) {
    const int nb = m->neigh[MESH_MAXNEIGH * t + k];
    if (nb < 0) return;
    check_node_recursive(m, nb, p, c);
  }
}

static void handle_corner(orc_mesh *m, const orc_params *prm, int t, int a, corner_scan *c) {
  const double *p = &m->node[9 * t + 3 * a];
  c->nvisited = c->nhas = 0;
  c->anyActive = 0;
  check_node_recursive(m, t, p, c);
  /* Domain::is_in_subdomain, domain_I.h:65-86 (serial: subdomain = box, padded) */
  for (int d = 0; d < 3; d++) {
    const double lo = prm->boxlo[d] - SMALL_DMBRDR, hi = prm->boxhi[d] + SMALL_DMBRDR;
    if (!(p[d] >= lo && p[d] < hi)) return;
  }
  int maxId = -1;
  for (int k = 0; k < c->nhas; k++)
    if (c->has[k] > maxId) maxId = c->has[k];
  int colinear = 0;
  const int ne = 2 * c->nhas;
  for (int i = 0; i < ne; i++)
    for (int j = i + 1; j < ne; j++)
      if (fabs(v3_dot(&c->edges[3 * i], &c->edges[3 * j])) > m->curvature &&
          !nodes_equal(m, &c->ends[3 * i], &c->ends[3 * j]))
        colinear = 1;
  if (colinear || !c->anyActive)
    m->cact[3 * t + a] = 0;
  else
    m->cact[3 * t + a] = (t == maxId) ? 1 : 0;
}

static void mesh_topology(orc_mesh *m, const orc_params *prm) {
  const int T = m->ntri;
  for (int t = 0; t < T; t++) {
    m->nneigh[t] = 0;
    for (int k = 0; k < MESH_MAXNEIGH; k++) m->neigh[MESH_MAXNEIGH * t + k] = -1;
    for (int k = 0; k < 3; k++) m->eact[3 * t + k] = m->cact[3 * t + k] = 1;
  }
  /* every pair sharing an edge, each handled once with the later element as i */
  for (int i = 0; i < T; i++)
    for (int j = 0; j < i; j++) {
      int ie, je;
      if (share_edge(m, i, j, &ie, &je)) {
        const int coplanar = fabs(v3_dot(&m->snorm[3 * i], &m->snorm[3 * j])) >= m->curvature;
        handle_shared_edge(m, i, ie, j, je, coplanar);
      }
    }
  corner_scan c;
  c.visited = (int *)xcalloc((size_t)T, sizeof(int));
  c.has = (int *)xcalloc((size_t)T, sizeof(int));
  c.edges = (double *)xcalloc(6 * (size_t)T, sizeof(double));
  c.ends = (double *)xcalloc(6 * (size_t)T, sizeof(double));
  for (int t = 0; t < T; t++)
    for (int a = 0; a < 3; a++) handle_corner(m, prm, t, a, &c);
  free(c.visited);
  free(c.has);
  free(c.edges);
  free(c.ends);
}

/* SurfaceMesh::areCoplanarNodeNeighs, surface_mesh_I.h:892-929 */
static int coplanar_node_neighs(const orc_mesh *m, int a, int b) {
  int areNeighs = 0;
  for (int k = 0; k < m->nneigh[a] && k < MESH_MAXNEIGH; k++)
    if (m->neigh[MESH_MAXNEIGH * a + k] == b) areNeighs = 1;
  int found = 0;
  if (!areNeighs && n_shared_nodes(m, a, b) != 0) found = 1;
  if (!areNeighs && !found) return 0;
  return fabs(v3_dot(&m->snorm[3 * a], &m->snorm[3 * b])) > m->curvature;
}

static void mesh_free(orc_mesh *m) {
  free(m->node);
  free(m->center);
  free(m->rbound);
  free(m->edgevec);
  free(m->edgelen);
  free(m->snorm);
  free(m->edgenorm);
  free(m->vnode);
  free(m->node_last);
  free(m->obtuse);
  free(m->eact);
  free(m->cact);
  free(m->nneigh);
  free(m->neigh);
  free(m->ncand);
  free(m->cand);
  free(m->npartner);
  free(m->ptri);
  free(m->keep);
  free(m->phist);
}

int orc_add_mesh(orc_sys *s, int ntri, const double *nodes, int wall_type,
                 const orc_model *model) {
  if (s->nmesh >= ORC_MAXMESH || ntri <= 0) return -1;
  orc_mesh *m = &s->mesh[s->nmesh];
  memset(m, 0, sizeof(*m));
  m->curvature = 1. - EPSILON_CURVATURE;
  m->precision = EPSILON_PRECISION;
  m->wall_type = wall_type;
  m->model = *model;
  const size_t T = (size_t)ntri;
  m->node = (double *)xcalloc(9 * T, sizeof(double));
  m->center = (double *)xcalloc(3 * T, sizeof(double));
  m->rbound = (double *)xcalloc(T, sizeof(double));
  m->edgevec = (double *)xcalloc(9 * T, sizeof(double));
  m->edgelen = (double *)xcalloc(3 * T, sizeof(double));
  m->snorm = (double *)xcalloc(3 * T, sizeof(double));
  m->edgenorm = (double *)xcalloc(9 * T, sizeof(double));
  m->vnode = (double *)xcalloc(9 * T, sizeof(double));
  m->node_last = (double *)xcalloc(9 * T, sizeof(double));
  m->obtuse = (int *)xcalloc(T, sizeof(int));
  m->eact = (unsigned char *)xcalloc(3 * T, 1);
  m->cact = (unsigned char *)xcalloc(3 * T, 1);
  m->nneigh = (int *)xcalloc(T, sizeof(int));
  m->neigh = (int *)xcalloc(MESH_MAXNEIGH * T, sizeof(int));
  /* MultiNodeMesh::addElement drops elements with coinciding nodes */
  int nt = 0;
  for (int t = 0; t < ntri; t++) {
    const double *q = &nodes[9 * t];
    if (nodes_equal(m, &q[0], &q[3]) || nodes_equal(m, &q[3], &q[6]) || nodes_equal(m, &q[0], &q[6]))
      continue;
    memcpy(&m->node[9 * nt], q, 9 * sizeof(double));
    double avg[3] = {0., 0., 0.};
    for (int i = 0; i < 3; i++)
      for (int k = 0; k < 3; k++) avg[k] = q[3 * i + k] + avg[k];
    v3_div(avg, 3.0);
    memcpy(&m->center[3 * nt], avg, 3 * sizeof(double));
    double rb = 0.;
    for (int i = 0; i < 3; i++) {
      double v[3];
      v3_sub(avg, &q[3 * i], v);
      const double mg = v3_mag(v);
      rb = rb > mg ? rb : mg;
    }
    m->rbound[nt] = rb;
    mesh_surf_props(m, nt);
    nt++;
  }
  m->ntri = nt;
  mesh_topology(m, &s->p);
  const size_t n = (size_t)(s->n > 0 ? s->n : 1);
  m->ncand = (int *)xcalloc(n, sizeof(int));
  m->cand = (int *)xcalloc(n * MESH_MAXCAND, sizeof(int));
  m->npartner = (int *)xcalloc(n, sizeof(int));
  m->ptri = (int *)xcalloc(n * MESH_MAXSLOT, sizeof(int));
  for (size_t k = 0; k < n * MESH_MAXSLOT; k++) m->ptri[k] = -1;
  m->keep = (unsigned char *)xcalloc(n * MESH_MAXSLOT, 1);
  m->phist = (double *)xcalloc(n * MESH_MAXSLOT * MESH_HW, sizeof(double));
  return s->nmesh++;
}

void orc_mesh_rotate(orc_sys *s, int imesh, const double *origin, const double *axis,
                     double period) {
  orc_mesh *m = &s->mesh[imesh];
  m->moving = 1;
  memcpy(m->origin, origin, 3 * sizeof(double));
  /* MeshMoverRotate: vectorNormalize3D(axis_), omega_ = 2 pi / period */
  const double norm = sqrt(axis[0] * axis[0] + axis[1] * axis[1] + axis[2] * axis[2]);
  const double invnorm = (norm == 0.) ? 0. : 1. / norm;
  for (int k = 0; k < 3; k++) m->axis[k] = axis[k] * invnorm;
  m->omega = 2. * MY_PI / period; /* M_PI in the reference: same constant */
}

/* fix mesh/surface/stress reference_point (mesh_module_stress.cpp:80-96) */
void orc_mesh_stress(orc_sys *s, int imesh, const double *p_ref) {
  orc_mesh *m = &s->mesh[imesh];
  m->stress = 1;
  for (int k = 0; k < 3; k++) m->p_ref[k] = p_ref ? p_ref[k] : 0.0;
}

void orc_mesh_force(const orc_sys *s, int imesh, double *force, double *torque) {
  const orc_mesh *m = &s->mesh[imesh];
  memcpy(force, m->ftot, sizeof(m->ftot));
  memcpy(torque, m->ttot, sizeof(m->ttot));
}

/* MeshMoverLinear, mesh_mover_linear.cpp:62-75 */
void orc_mesh_linear(orc_sys *s, int imesh, const double *vel) {
  orc_mesh *m = &s->mesh[imesh];
  m->moving = 2;
  memcpy(m->vel, vel, 3 * sizeof(double));
}

/* MeshMoverVibLin (mesh_mover_linear.cpp:230-263) and MeshMoverVibRot
 * (mesh_mover_rotation.cpp:333-375): unit axis, omega_j = 2 pi / period_j */
static void set_vib(orc_mesh *m, const double *axis, int order, const double *amplitude,
                    const double *phase, const double *period) {
  const double norm = sqrt(axis[0] * axis[0] + axis[1] * axis[1] + axis[2] * axis[2]);
  const double invnorm = (norm == 0.) ? 0. : 1. / norm;
  for (int k = 0; k < 3; k++) m->axis[k] = axis[k] * invnorm;
  m->vord = order;
  for (int j = 0; j < order; j++) {
    m->vampl[j] = amplitude[j];
    m->vphi[j] = phase[j];
    m->vomega[j] = 2. * MY_PI / period[j];
  }
}

void orc_mesh_viblin(orc_sys *s, int imesh, const double *axis, int order, const double *amplitude,
                     const double *phase, const double *period) {
  orc_mesh *m = &s->mesh[imesh];
  m->moving = 4;
  set_vib(m, axis, order, amplitude, phase, period);
}

void orc_mesh_vibrot(orc_sys *s, int imesh, const double *origin, const double *axis, int order,
                     const double *amplitude, const double *phase, const double *period) {
  orc_mesh *m = &s->mesh[imesh];
  m->moving = 5;
  memcpy(m->origin, origin, 3 * sizeof(double));
  set_vib(m, axis, order, amplitude, phase, period);
}

/* MeshMoverWiggle, mesh_mover_linear.cpp:160-183 */
void orc_mesh_wiggle(orc_sys *s, int imesh, const double *amplitude, double period) {
  orc_mesh *m = &s->mesh[imesh];
  m->moving = 3;
  memcpy(m->ampl, amplitude, 3 * sizeof(double));
  m->womega = 2. * MY_PI / period;
}

void orc_mesh_topology(const orc_sys *s, int imesh, int *edge_active, int *corner_active,
                       int *obtuse) {
  const orc_mesh *m = &s->mesh[imesh];
  for (int t = 0; t < m->ntri; t++) {
    for (int k = 0; k < 3; k++) {
      edge_active[3 * t + k] = m->eact[3 * t + k];
      corner_active[3 * t + k] = m->cact[3 * t + k];
    }
    obtuse[t] = m->obtuse[t];
  }
}

void orc_mesh_nodes(const orc_sys *s, int imesh, double *nodes, double *vnodes) {
  const orc_mesh *m = &s->mesh[imesh];
  if (nodes) memcpy(nodes, m->node, 9 * sizeof(double) * (size_t)m->ntri);
  if (vnodes) memcpy(vnodes, m->vnode, 9 * sizeof(double) * (size_t)m->ntri);
}

int orc_mesh_history(const orc_sys *s, int imesh, int maxslot, int *tri, double *hist) {
  const orc_mesh *m = &s->mesh[imesh];
  for (int i = 0; i < s->n; i++)
    for (int k = 0; k < maxslot; k++) {
      const int inside = k < MESH_MAXSLOT;
      const int tt = inside ? m->ptri[(size_t)i * MESH_MAXSLOT + k] : -1;
      tri[(size_t)i * maxslot + k] = tt;
      for (int d = 0; d < 3; d++)
        hist[((size_t)i * maxslot + k) * 3 + d] =
            (inside && tt >= 0) ? m->phist[((size_t)i * MESH_MAXSLOT + k) * MESH_HW + d] : 0.0;
    }
  return MESH_MAXSLOT;
}

int64_t orc_mesh_forced_rebuilds(const orc_sys *s) { return s->mesh_forced; }

/* ---- sphere-triangle contact: TriMesh::resolveTriSphereContactBary ---- */
static double calc_dist(const double *p, const double *cp, double *delta) {
  v3_sub(cp, p, delta);
  return sqrt((p[0] - cp[0]) * (p[0] - cp[0]) + (p[1] - cp[1]) * (p[1] - cp[1]) +
              (p[2] - cp[2]) * (p[2] - cp[2]));
}

static double resolve_edge(const orc_mesh *m, int t, int iEdge, const double *p, double *delta,
                           double *bary) {
  const int ip = (iEdge + 1) % 3, ipp = (iEdge + 2) % 3;
  const double *n = &m->node[9 * t];
  const double *ev = &m->edgevec[9 * t + 3 * iEdge];
  const double len = m->edgelen[3 * t + iEdge];
  double nodeToP[3];
  v3_sub(p, &n[3 * iEdge], nodeToP);
  const double distFromNode = v3_dot(nodeToP, ev);
  if (distFromNode < -SMALL_TRIMESH) {
    if (!m->cact[3 * t + iEdge]) return LARGE_TRIMESH;
    const double d = calc_dist(p, &n[3 * iEdge], delta);
    bary[iEdge] = 1.;
    bary[ip] = 0.;
    bary[ipp] = 0.;
    return d;
  } else if (distFromNode > len + SMALL_TRIMESH) {
    if (!m->cact[3 * t + ip]) return LARGE_TRIMESH;
    const double d = calc_dist(p, &n[3 * ip], delta);
    bary[iEdge] = 0.;
    bary[ip] = 1.;
    bary[ipp] = 0.;
    return d;
  } else {
    if (!m->eact[3 * t + iEdge]) return LARGE_TRIMESH;
    double cp[3];
    for (int k = 0; k < 3; k++) cp[k] = n[3 * iEdge + k] + distFromNode * ev[k];
    const double d = calc_dist(p, cp, delta);
    bary[ipp] = 0.;
    bary[iEdge] = 1. - distFromNode / len;
    bary[ip] = 1. - bary[iEdge];
    return d;
  }
}

static double resolve_corner(const orc_mesh *m, int t, int iNode, int obtuse, const double *p,
                             double *delta, double *bary) {
  const int ip = (iNode + 1) % 3, ipp = (iNode + 2) % 3;
  const double *n = &m->node[9 * t + 3 * iNode];
  if (obtuse) {
    const double *edge = &m->edgevec[9 * t];
    const double *len = &m->edgelen[3 * t];
    double nodeToP[3], cp[3];
    v3_sub(p, n, nodeToP);
    double distFromNode = v3_dot(nodeToP, &edge[3 * ipp]);
    if (distFromNode < SMALL_TRIMESH) {
      if (distFromNode > -len[ipp]) {
        if (!m->eact[3 * t + ipp]) return LARGE_TRIMESH;
        for (int k = 0; k < 3; k++) cp[k] = n[k] + distFromNode * edge[3 * ipp + k];
        bary[ip] = 0.;
        bary[iNode] = 1. + distFromNode / len[ipp];
        bary[ipp] = 1. - bary[iNode];
        return calc_dist(p, cp, delta);
      } else {
        if (!m->cact[3 * t + ipp]) return LARGE_TRIMESH;
        bary[ipp] = 1.;
        bary[iNode] = bary[ip] = 0.;
        return calc_dist(p, &m->node[9 * t + 3 * ipp], delta);
      }
    }
    distFromNode = v3_dot(nodeToP, &edge[3 * iNode]);
    if (distFromNode > -SMALL_TRIMESH) {
      if (distFromNode < len[iNode]) {
        if (!m->eact[3 * t + iNode]) return LARGE_TRIMESH;
        for (int k = 0; k < 3; k++) cp[k] = n[k] + distFromNode * edge[3 * iNode + k];
        bary[ipp] = 0.;
        bary[iNode] = 1. - distFromNode / len[iNode];
        bary[ip] = 1. - bary[iNode];
        return calc_dist(p, cp, delta);
      } else {
        if (!m->cact[3 * t + ip]) return LARGE_TRIMESH;
        bary[ip] = 1.;
        bary[iNode] = bary[ipp] = 0.;
        return calc_dist(p, &m->node[9 * t + 3 * ip], delta);
      }
    }
  }
  if (!m->cact[3 * t + iNode]) return LARGE_TRIMESH;
  bary[iNode] = 1.;
  bary[ip] = bary[ipp] = 0.;
  return calc_dist(p, n, delta);
}

static double resolve_tri_sphere(const orc_mesh *m, int t, double r, const double *p,
                                 double *delta, double *bary, int *barysign) {
  const double *n = &m->node[9 * t];
  const int obt = m->obtuse[t];
  bary[0] = bary[1] = bary[2] = 0.;
  double n0p[3];
  v3_sub(p, n, n0p);
  /* MathExtraLiggghts::calcBaryTriCoords */
  const double *ev = &m->edgevec[9 * t];
  const double *el = &m->edgelen[3 * t];
  const double a = v3_dot(n0p, &ev[0]);
  const double b = v3_dot(n0p, &ev[6]);
  const double c = v3_dot(&ev[0], &ev[6]);
  const double oneMinCSqr = 1 - c * c;
  bary[1] = (a - b * c) / (el[0] * oneMinCSqr);
  bary[2] = (a * c - b) / (el[2] * oneMinCSqr);
  bary[0] = 1. - bary[1] - bary[2];
  const double invlen = 1. / (2. * m->rbound[t]);
  const int bs = (bary[0] > -m->precision * invlen) + 2 * (bary[1] > -m->precision * invlen) +
                 4 * (bary[2] > -m->precision * invlen);
  *barysign = bs;
  double d = 0.;
  switch (bs) {
    case 1: d = resolve_corner(m, t, 0, obt == 0, p, delta, bary); break;
    case 2: d = resolve_corner(m, t, 1, obt == 1, p, delta, bary); break;
    case 3: d = resolve_edge(m, t, 0, p, delta, bary); break;
    case 4: d = resolve_corner(m, t, 2, obt == 2, p, delta, bary); break;
    case 5: d = resolve_edge(m, t, 2, p, delta, bary); break;
    case 6: d = resolve_edge(m, t, 1, p, delta, bary); break;
    case 7: {
      const double *sn = &m->snorm[3 * t];
      const double dNorm = v3_dot(sn, n0p);
      double cs[3];
      for (int k = 0; k < 3; k++) cs[k] = p[k] - sn[k] * dNorm;
      d = calc_dist(p, cs, delta);
      break;
    }
    default: d = 1.; break; /* all three coordinates negative: impossible */
  }
  return d - r;
}

double orc_tri_contact(const double *nodes9, const int *edge_active3, const int *corner_active3,
                       const double *p, double radius, double *delta, double *bary,
                       int *barysign) {
  orc_mesh m;
  memset(&m, 0, sizeof(m));
  double node[9], center[3], rb, ev[9], el[3], sn[3], en[9];
  int obt;
  unsigned char ea[3], ca[3];
  memcpy(node, nodes9, sizeof node);
  m.node = node;
  m.center = center;
  m.rbound = &rb;
  m.edgevec = ev;
  m.edgelen = el;
  m.snorm = sn;
  m.edgenorm = en;
  m.obtuse = &obt;
  m.eact = ea;
  m.cact = ca;
  m.precision = EPSILON_PRECISION;
  double avg[3] = {0., 0., 0.};
  for (int i = 0; i < 3; i++)
    for (int k = 0; k < 3; k++) avg[k] = node[3 * i + k] + avg[k];
  v3_div(avg, 3.0);
  memcpy(center, avg, sizeof avg);
  rb = 0.;
  for (int i = 0; i < 3; i++) {
    double v[3];
    v3_sub(avg, &node[3 * i], v);
    const double mg = v3_mag(v);
    rb = rb > mg ? rb : mg;
  }
  mesh_surf_props(&m, 0);
  for (int k = 0; k < 3; k++) {
    ea[k] = (unsigned char)(edge_active3[k] != 0);
    ca[k] = (unsigned char)(corner_active3[k] != 0);
  }
  return resolve_tri_sphere(&m, 0, radius, p, delta, bary, barysign);
}

/* closest point on a triangle to p (Ericson, Real-Time Collision Detection
 * 5.1.5); used only to select candidate triangles within r + skin */
static double tri_point_dist2(const double *a, const double *b, const double *c, const double *p) {
  double ab[3], ac[3], ap[3], q[3];
  v3_sub(b, a, ab);
  v3_sub(c, a, ac);
  v3_sub(p, a, ap);
  const double d1 = v3_dot(ab, ap), d2 = v3_dot(ac, ap);
  if (d1 <= 0. && d2 <= 0.) {
    memcpy(q, a, sizeof q);
  } else {
    double bp[3];
    v3_sub(p, b, bp);
    const double d3 = v3_dot(ab, bp), d4 = v3_dot(ac, bp);
    if (d3 >= 0. && d4 <= d3) {
      memcpy(q, b, sizeof q);
    } else {
      const double vc = d1 * d4 - d3 * d2;
      if (vc <= 0. && d1 >= 0. && d3 <= 0.) {
        const double v = d1 / (d1 - d3);
        for (int k = 0; k < 3; k++) q[k] = a[k] + v * ab[k];
      } else {
        double cp[3];
        v3_sub(p, c, cp);
        const double d5 = v3_dot(ab, cp), d6 = v3_dot(ac, cp);
        if (d6 >= 0. && d5 <= d6) {
          memcpy(q, c, sizeof q);
        } else {
          const double vb = d5 * d2 - d1 * d6;
          if (vb <= 0. && d2 >= 0. && d6 <= 0.) {
            const double w = d2 / (d2 - d6);
            for (int k = 0; k < 3; k++) q[k] = a[k] + w * ac[k];
          } else {
            const double va = d3 * d6 - d5 * d4;
            if (va <= 0. && (d4 - d3) >= 0. && (d5 - d6) >= 0.) {
              const double w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
              for (int k = 0; k < 3; k++) q[k] = b[k] + w * (c[k] - b[k]);
            } else {
              const double denom = 1. / (va + vb + vc);
              const double v = vb * denom, w = vc * denom;
              for (int k = 0; k < 3; k++) q[k] = a[k] + ab[k] * v + ac[k] * w;
            }
          }
        }
      }
    }
  }
  double d[3];
  v3_sub(p, q, d);
  return v3_dot(d, d);
}

/* per-particle slot arrays */
#define SLOT(m, i, k) ((m)->ptri[(size_t)(i) * MESH_MAXSLOT + (k)])
#define SKEEP(m, i, k) ((m)->keep[(size_t)(i) * MESH_MAXSLOT + (k)])
#define SHIST(m, i, k) (&(m)->phist[((size_t)(i) * MESH_MAXSLOT + (k)) * MESH_HW])

static void slot_swap(orc_mesh *m, int i, int a, int b, int keepflag) {
  int t = SLOT(m, i, a);
  SLOT(m, i, a) = SLOT(m, i, b);
  SLOT(m, i, b) = t;
  double h[MESH_HW];
  memcpy(h, SHIST(m, i, a), sizeof h);
  memcpy(SHIST(m, i, a), SHIST(m, i, b), sizeof h);
  memcpy(SHIST(m, i, b), h, sizeof h);
  if (keepflag) {
    const unsigned char kf = SKEEP(m, i, a);
    SKEEP(m, i, a) = SKEEP(m, i, b);
    SKEEP(m, i, b) = kf;
  }
}

/* mesh part of a neighbor rebuild: FixContactHistoryMesh::sort_contacts
 * (pre_exchange), element geometry refresh and node positions at rebuild
 * (FixMesh::pre_force -> pbcExchangeBorders -> refreshOwned), candidate
 * triangles within r + skin, FixContactHistoryMesh::cleanUpContactJumps */
static void mesh_rebuild(orc_sys *s) {
  for (int im = 0; im < s->nmesh; im++) {
    orc_mesh *m = &s->mesh[im];
    for (int i = 0; i < s->n; i++) {
      const int nslot = MESH_MAXSLOT;
      int first_empty, last_filled;
      do {
        first_empty = last_filled = -1;
        for (int j = 0; j < nslot; j++) {
          if (-1 == first_empty && -1 == SLOT(m, i, j)) first_empty = j;
          if (SLOT(m, i, j) >= 0) last_filled = j;
        }
        if (first_empty > -1 && last_filled > -1 && first_empty < last_filled)
          slot_swap(m, i, first_empty, last_filled, 1);
      } while (first_empty > -1 && last_filled > -1 && first_empty < last_filled);
    }
    for (int t = 0; t < m->ntri; t++) mesh_surf_props(m, t);
    memcpy(m->node_last, m->node, 9 * sizeof(double) * (size_t)m->ntri);
    for (int i = 0; i < s->n; i++) {
      const double lim = s->r[i] * s->p.cdf + s->p.skin;
      int nc = 0;
      for (int t = 0; t < m->ntri; t++) {
        const double *q = &m->node[9 * t];
        if (tri_point_dist2(&q[0], &q[3], &q[6], &s->x[3 * i]) <= lim * lim) {
          if (nc < MESH_MAXCAND) m->cand[(size_t)i * MESH_MAXCAND + nc] = t;
          nc++;
        }
      }
      if (nc > MESH_MAXCAND) abort();
      m->ncand[i] = nc;
      /* cleanUpContactJumps */
      int ip = 0;
      while (ip < m->npartner[i]) {
        const int tt = SLOT(m, i, ip);
        int inlist = 0;
        for (int k = 0; k < nc; k++)
          if (m->cand[(size_t)i * MESH_MAXCAND + k] == tt) inlist = 1;
        if (!inlist) {
          SLOT(m, i, ip) = -1;
          memset(SHIST(m, i, ip), 0, MESH_HW * sizeof(double));
          slot_swap(m, i, ip, m->npartner[i] - 1, 0);
          m->npartner[i]--;
        } else {
          ip++;
        }
      }
    }
  }
}

/* FixMoveMesh::initial_integrate (fix_move_mesh.cpp:221-238) -> the mover's
 * initial_integrate: MeshMoverRotate (mesh_mover_rotation.cpp:110-139),
 * MeshMoverLinear / MeshMoverWiggle (mesh_mover_linear.cpp:88-101, 203-218) */
static void mesh_move(orc_sys *s) {
  for (int im = 0; im < s->nmesh; im++) {
    orc_mesh *m = &s->mesh[im];
    if (!m->moving) continue;
    m->time += s->p.dt; /* time_ += update->dt */
    /* viblin / vibrot: vA = sum_j -a_j w_j sin(w_j t + phi_j) */
    double vA = 0.0;
    if (m->moving >= 4)
      for (int j = 0; j < m->vord; j++)
        vA = vA - m->vampl[j] * m->vomega[j] * sin(m->vomega[j] * m->time + m->vphi[j]);
    if (m->moving != 1 && m->moving != 5) {
      /* translation: v_node reset to zero, vNode = vel_ or
       * amplitude_ * (omega_ cos(omega_ time_)), dx = vNode dt; nodes and
       * centre += dx (MultiNodeMesh::move, multi_node_mesh_I.h:502-527) */
      double vN[3], dx[3];
      if (m->moving == 3) {
        const double sc = m->womega * cos(m->womega * m->time);
        for (int k = 0; k < 3; k++) vN[k] = m->ampl[k] * sc;
      } else if (m->moving == 4) {
        for (int k = 0; k < 3; k++) vN[k] = m->axis[k] * vA;
      } else {
        for (int k = 0; k < 3; k++) vN[k] = m->vel[k];
      }
      for (int k = 0; k < 3; k++) dx[k] = vN[k] * s->p.dt;
      for (int t = 0; t < m->ntri; t++) {
        for (int j = 0; j < 3; j++)
          for (int k = 0; k < 3; k++) m->node[9 * t + 3 * j + k] += dx[k];
        for (int k = 0; k < 3; k++) m->center[3 * t + k] += dx[k];
        for (int j = 0; j < 3; j++)
          for (int k = 0; k < 3; k++) m->vnode[9 * t + 3 * j + k] = 0.0 + vN[k];
      }
      continue;
    }
    /* rotate: omega dt; vibrot: vR dt (mesh_mover_rotation.cpp:397-430) */
    const double dAngle = (m->moving == 5 ? vA : m->omega) * s->p.dt;
    double axisNorm[3] = {m->axis[0], m->axis[1], m->axis[2]};
    v3_div(axisNorm, v3_mag(axisNorm));
    double dQ[4];
    dQ[0] = cos(dAngle * 0.5);
    for (int k = 0; k < 3; k++) dQ[k + 1] = axisNorm[k] * sin(dAngle * 0.5);
    const double *o = m->origin;
    const int trans = (o[0] * o[0] + o[1] * o[1] + o[2] * o[2]) > 0.;
    for (int t = 0; t < m->ntri; t++) {
      double *c = &m->center[3 * t];
      c[0] = c[1] = c[2] = 0.;
      for (int j = 0; j < 3; j++) {
        double *nd = &m->node[9 * t + 3 * j];
        if (trans) v3_sub(nd, o, nd);
        vec_quat_rotate(nd, dQ);
        if (trans)
          for (int k = 0; k < 3; k++) nd[k] = nd[k] + o[k];
        for (int k = 0; k < 3; k++) c[k] = nd[k] + c[k];
      }
      v3_div(c, 3.0);
      /* edgeVec, edgeNorm, surfaceNorm are rotated with the mesh */
      for (int j = 0; j < 3; j++) {
        vec_quat_rotate(&m->edgevec[9 * t + 3 * j], dQ);
        vec_quat_rotate(&m->edgenorm[9 * t + 3 * j], dQ);
      }
      vec_quat_rotate(&m->snorm[3 * t], dQ);
    }
    /* node velocity: reset, then v += (omega axis) x (node - origin) */
    double omegaVec[3];
    for (int k = 0; k < 3; k++) omegaVec[k] = m->axis[k] * (m->moving == 5 ? vA : m->omega);
    for (int t = 0; t < m->ntri; t++)
      for (int j = 0; j < 3; j++) {
        double rPA[3], vRot[3];
        v3_sub(&m->node[9 * t + 3 * j], o, rPA);
        v3_cross(omegaVec, rPA, vRot);
        double *v = &m->vnode[9 * t + 3 * j];
        for (int k = 0; k < 3; k++) v[k] = 0.0 + vRot[k];
      }
  }
}

/* MultiNodeMesh::decideRebuild */
static int mesh_decide_rebuild(const orc_sys *s) {
  const double triggersq = 0.25 * s->p.skin * s->p.skin;
  for (int im = 0; im < s->nmesh; im++) {
    const orc_mesh *m = &s->mesh[im];
    if (!m->moving) continue;
    for (int k = 0; k < 3 * m->ntri; k++) {
      const double *a = &m->node[3 * k], *b = &m->node_last[3 * k];
      const double dx = a[0] - b[0], dy = a[1] - b[1], dz = a[2] - b[2];
      if (dx * dx + dy * dy + dz * dz > triggersq) return 1;
    }
  }
  return 0;
}

/* FixContactHistoryMesh::handleContact */
static int mesh_handle_contact(orc_mesh *m, int i, int tri, double **hist, int faceflag) {
  for (int k = 0; k < MESH_MAXSLOT; k++)
    if (SLOT(m, i, k) == tri) {
      *hist = SHIST(m, i, k);
      SKEEP(m, i, k) = 1;
      return 1;
    }
  if (faceflag) {
    for (int k = 0; k < MESH_MAXSLOT; k++) {
      const int pt = SLOT(m, i, k);
      if (pt >= 0 && pt != tri && coplanar_node_neighs(m, pt, tri) && SKEEP(m, i, k)) return 0;
    }
  }
  int ic = -1;
  for (int k = 0; k < MESH_MAXSLOT; k++)
    if (SLOT(m, i, k) == -1) {
      ic = k;
      break;
    }
  if (ic < 0) abort(); /* more simultaneous mesh contacts than slots */
  SLOT(m, i, ic) = tri;
  SKEEP(m, i, ic) = 1;
  m->npartner[i]++;
  *hist = SHIST(m, i, ic);
  memset(*hist, 0, MESH_HW * sizeof(double));
  if (faceflag)
    for (int k = 0; k < MESH_MAXSLOT; k++) {
      const int pt = SLOT(m, i, k);
      if (pt >= 0 && pt != tri && coplanar_node_neighs(m, pt, tri))
        memcpy(*hist, SHIST(m, i, k), MESH_HW * sizeof(double));
    }
  return 1;
}

/* FixWallGran::post_force_mesh for one mesh */
static void mesh_wall_compute(orc_sys *s, int im, int shearupdate) {
  orc_mesh *m = &s->mesh[im];
  /* MeshModuleStress::pre_force (mesh_module_stress.cpp:225-236): totals restart */
  memset(m->ftot, 0, sizeof(m->ftot));
  memset(m->ttot, 0, sizeof(m->ttot));
  for (int i = 0; i < s->n; i++) {
    for (int k = 0; k < MESH_MAXSLOT; k++) SKEEP(m, i, k) = 0; /* markAllContacts */
    for (int c = 0; c < m->ncand[i]; c++) {
      const int t = m->cand[(size_t)i * MESH_MAXCAND + c];
      double delta[3], bary[3];
      int barysign;
      const double deltan = resolve_tri_sphere(m, t, s->r[i], &s->x[3 * i], delta, bary, &barysign);
      if (deltan > s->cutneighmax) continue;
      if (deltan <= 0 || deltan < (s->p.cdf - 1.0) * s->r[i]) {
        double *h = NULL;
        if (!mesh_handle_contact(m, i, t, &h, 7 == barysign)) continue;
        double vwall[3] = {0., 0., 0.};
        if (m->moving)
          for (int k = 0; k < 3; k++)
            vwall[k] = (bary[0] * m->vnode[9 * t + k] + bary[1] * m->vnode[9 * t + 3 + k] +
                        bary[2] * m->vnode[9 * t + 6 + k]);
        if (!(deltan <= 0)) {
          /* compute_force with intersectflag false (fix_wall_gran.cpp:987-1006,
           * fix_wall_gran_base.h:283-289): surfacesClose zeroes the slot's shear
           * (tangential_model_history.h:358-369) and the epsd2 rolling torque
           * (rolling_model_epsd2.h:144-151); no force */
          if (m->model.tangential_model == 0) h[0] = h[1] = h[2] = 0.0;
          if (m->model.rolling_epsd2) h[3] = h[4] = h[5] = 0.0;
          continue;
        }
        const double sdelta[3] = {-delta[0], -delta[1], -delta[2]};
        double dF[3], dT[3];
        orc_contact_wall_r(&m->model, &s->m, s->type[i], m->wall_type, -deltan, sdelta, s->r[i],
                           &s->v[3 * i], &s->w[3 * i], vwall, s->rm[i], s->p.dt, shearupdate, h, h + 3,
                           dF, dT);
        double fold[3];
        for (int d = 0; d < 3; d++) {
          fold[d] = s->f[3 * i + d];
          s->f[3 * i + d] += dF[d];
          s->t[3 * i + d] += dT[d];
        }
        if (m->stress) {
          /* fix_wall_gran_base.h:344-359: f_pw = f - force_old, handed with
           * delta = -sidata.delta to MeshModuleStress::add_particle_contribution
           * (mesh_module_stress.cpp:276-300), which negates it and adds it and
           * its torque about p_ref at the contact point x + delta */
          double frc[3], cp[3], rr[3];
          for (int d = 0; d < 3; d++) frc[d] = -(s->f[3 * i + d] - fold[d]);
          for (int d = 0; d < 3; d++) cp[d] = s->x[3 * i + d] + (-sdelta[d]);
          for (int d = 0; d < 3; d++) {
            m->ftot[d] += frc[d];
            rr[d] = cp[d] - m->p_ref[d];
          }
          m->ttot[0] += rr[1] * frc[2] - rr[2] * frc[1];
          m->ttot[1] += rr[2] * frc[0] - rr[0] * frc[2];
          m->ttot[2] += rr[0] * frc[1] - rr[1] * frc[0];
        }
      }
    }
    /* cleanUpContacts */
    for (int k = 0; k < MESH_MAXSLOT; k++)
      if (!SKEEP(m, i, k)) {
        if (SLOT(m, i, k) > -1) m->npartner[i]--;
        SLOT(m, i, k) = -1;
        memset(SHIST(m, i, k), 0, MESH_HW * sizeof(double));
      }
  }
}

/* compute wall/gran/local (compute_pair_gran_local.cpp:648-870) through
 * FixWallGran::post_force_pgl (fix_wall_gran.cpp:752-765: computeflag 0,
 * shearupdate 0): the contact pass of mesh_wall_compute on private copies of
 * each particle's slot table and histories, so nothing changes. A row per
 * evaluated contact (touching, or in the close band, and accepted by
 * handleContact), mesh by mesh and particle by particle: ids = mesh, triangle,
 * tag; data = contact point x + delta[3], x[3], wall velocity at the contact[3],
 * v[3], force[3] and torque[3] on the particle (zero in the close band), unit
 * normal delta / r[3] (minus the contact's en, fix_wall_gran_base.h:332-334),
 * r^2 with r = radius + deltan (sidata.rsq, fix_wall_gran_base.h:181-182),
 * the radius, shear history[3]. Returns the number of rows (negated when above cap). */
int64_t orc_wall_local(const orc_sys *s, int64_t cap, int *ids, double *data) {
  enum { NC = 26 };
  int64_t nr = 0;
  for (int im = 0; im < s->nmesh; im++) {
    const orc_mesh *m = &s->mesh[im];
    for (int i = 0; i < s->n; i++) {
      int slot[MESH_MAXSLOT];
      unsigned char keep[MESH_MAXSLOT];
      double hist[MESH_MAXSLOT][MESH_HW];
      for (int k = 0; k < MESH_MAXSLOT; k++) {
        slot[k] = SLOT(m, i, k);
        keep[k] = 0;
        memcpy(hist[k], SHIST(m, i, k), sizeof hist[k]);
      }
      for (int c = 0; c < m->ncand[i]; c++) {
        const int t = m->cand[(size_t)i * MESH_MAXCAND + c];
        double delta[3], bary[3];
        int barysign;
        const double deltan = resolve_tri_sphere(m, t, s->r[i], &s->x[3 * i], delta, bary, &barysign);
        if (deltan > s->cutneighmax) continue;
        if (!(deltan <= 0 || deltan < (s->p.cdf - 1.0) * s->r[i])) continue;
        /* mesh_handle_contact on the copies */
        int ic = -1;
        for (int k = 0; k < MESH_MAXSLOT; k++)
          if (slot[k] == t) {
            ic = k;
            break;
          }
        if (ic >= 0) {
          keep[ic] = 1;
        } else {
          int skip = 0;
          if (7 == barysign)
            for (int k = 0; k < MESH_MAXSLOT; k++)
              if (slot[k] >= 0 && slot[k] != t && coplanar_node_neighs(m, slot[k], t) && keep[k]) skip = 1;
          if (skip) continue;
          for (int k = 0; k < MESH_MAXSLOT; k++)
            if (slot[k] == -1) {
              ic = k;
              break;
            }
          if (ic < 0) abort();
          slot[ic] = t;
          keep[ic] = 1;
          memset(hist[ic], 0, sizeof hist[ic]);
          if (7 == barysign)
            for (int k = 0; k < MESH_MAXSLOT; k++)
              if (slot[k] >= 0 && slot[k] != t && coplanar_node_neighs(m, slot[k], t))
                memcpy(hist[ic], hist[k], sizeof hist[ic]);
        }
        double *h = hist[ic];
        double vwall[3] = {0., 0., 0.};
        if (m->moving)
          for (int k = 0; k < 3; k++)
            vwall[k] = (bary[0] * m->vnode[9 * t + k] + bary[1] * m->vnode[9 * t + 3 + k] +
                        bary[2] * m->vnode[9 * t + 6 + k]);
        double dF[3] = {0., 0., 0.}, dT[3] = {0., 0., 0.};
        if (!(deltan <= 0)) {
          if (m->model.tangential_model == 0) h[0] = h[1] = h[2] = 0.0;
          if (m->model.rolling_epsd2) h[3] = h[4] = h[5] = 0.0;
        } else {
          const double sdelta[3] = {-delta[0], -delta[1], -delta[2]};
          orc_contact_wall_r(&m->model, &s->m, s->type[i], m->wall_type, -deltan, sdelta, s->r[i],
                             &s->v[3 * i], &s->w[3 * i], vwall, s->rm[i], s->p.dt, 0, h, h + 3, dF, dT);
        }
        if (nr < cap) {
          double *row = data + nr * NC;
          const double rr = s->r[i] + deltan;
          for (int d = 0; d < 3; d++) {
            row[d] = s->x[3 * i + d] + delta[d];
            row[3 + d] = s->x[3 * i + d];
            row[6 + d] = vwall[d];
            row[9 + d] = s->v[3 * i + d];
            row[12 + d] = dF[d];
            row[15 + d] = dT[d];
            row[18 + d] = delta[d] * (1.0 / rr);
            row[23 + d] = h[d];
          }
          row[21] = rr * rr;
          row[22] = s->r[i];
          ids[3 * nr] = im;
          ids[3 * nr + 1] = t;
          ids[3 * nr + 2] = s->tag[i];
        }
        nr++;
      }
    }
  }
  return nr > cap ? -nr : nr;
}
