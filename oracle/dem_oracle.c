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
    }
}

/* ------------------------------------------------------------------------
 * shared model chain for pair and wall contacts:
 *   SurfaceModel<DEFAULT>::surfacesIntersect   surface_model_default.h:140-205
 *   NormalModel<HERTZ>::surfacesIntersect      normal_model_hertz.h:202-405
 *   TangentialModel<HISTORY>::surfacesIntersect tangential_model_history.h:130-356
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
                        double dt, int shearupdate, double *shear, double dFi[3],
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

  /* normal model: Hertz */
  const int it = c->itype, jt = c->jtype;
  const double reff = c->is_wall ? c->radi : (c->radi * c->radj / (c->radi + c->radj));
  const double meff = c->meff;
  const double sqrtval = sqrt(reff * deltan);
  const double Sn = 2. * mat->Yeff[it][jt] * sqrtval;
  const double St = 8. * mat->Geff[it][jt] * sqrtval;
  double kn = 4. / 3. * mat->Yeff[it][jt] * sqrtval;
  double kt = St;
  const double sqrtFiveOverSix = 0.91287092917527685576161630466800355658790782499663875;
  const double gamman = -2. * sqrtFiveOverSix * mat->betaeff[it][jt] * sqrt(Sn * meff);
  const double gammat =
      mdl->tangential_damping ? -2. * sqrtFiveOverSix * mat->betaeff[it][jt] * sqrt(St * meff)
                              : 0.0;
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
  const double xmu = mat->mu[it][jt];
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

  /* rolling model: constant directional torque */
  if (mdl->rolling_cdt) {
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
  model_chain(mdl, mat, &c, dt, shearupdate, shear, dFi, dTi, dTj);
  return 1;
}

/* Walls::Granular<CM>::compute_force, fix_wall_gran_base.h:148-361 */
void orc_contact_wall(const orc_model *mdl, const orc_materials *mat, int itype, int wtype,
                      double sidata_deltan, const double *sidata_delta, double radi,
                      const double *vi, const double *wi, const double *vwall, double mi,
                      double dt, int shearupdate, double *shear, double *dFi, double *dTi) {
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
  model_chain(mdl, mat, &c, dt, shearupdate, shear, dFi, dTi, dTj);
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
  int64_t cap;
  int built;
  /* per-atom partner arrays, fix_contact_history.cpp:305-425 */
  int *npartner;
  int64_t *poff;
  int *partner;
  double *phist;
  int64_t pcap;
  /* primitive wall history, one property/atom vector per wall (fix_wall_gran.cpp:498-522) */
  double *whist;
  int64_t ntimestep;
  int64_t nrebuilds;
  double cutneighmax;
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
  /* cutneighmax = max_ij (rmax_i + rmax_j) * cdf + skin, pair_gran.cpp:623-635, neighbor.cpp:310-332 */
  double rmax = 0.0;
  for (int i = 0; i < n; i++)
    if (radius[i] > rmax) rmax = radius[i];
  s->cutneighmax = (rmax + rmax) * p->cdf + p->skin;
  return s;
}

void orc_destroy(orc_sys *s) {
  if (!s) return;
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
  free(s->npartner);
  free(s->poff);
  free(s->partner);
  free(s->phist);
  free(s->whist);
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
    s->partner = (int *)xcalloc((size_t)s->pcap, sizeof(int));
    s->phist = (double *)xcalloc(3 * (size_t)s->pcap, sizeof(double));
  }
  memset(s->npartner, 0, sizeof(int) * (size_t)n);
  for (int i = 0; i < n; i++)
    for (int64_t e = s->first[i]; e < s->first[i + 1]; e++)
      if (s->flag[e]) {
        const int j = s->nbr[e];
        int64_t m = s->poff[i] + s->npartner[i]++;
        s->partner[m] = s->tag[j];
        for (int d = 0; d < 3; d++) s->phist[3 * m + d] = s->hist[3 * e + d];
        m = s->poff[j] + s->npartner[j]++;
        s->partner[m] = s->tag[i];
        for (int d = 0; d < 3; d++) s->phist[3 * m + d] = -s->hist[3 * e + d];
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
                int fl = 0;
                h[0] = h[1] = h[2] = 0.0;
                if (rsq < radsum * radsum) {
                  for (int64_t m = s->poff[i]; m < s->poff[i] + s->npartner[i]; m++)
                    if (s->partner[m] == s->tag[j]) {
                      fl = 1;
                      h[0] = s->phist[3 * m];
                      h[1] = s->phist[3 * m + 1];
                      h[2] = s->phist[3 * m + 2];
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
      s->nbr = (int *)xcalloc((size_t)s->cap, sizeof(int));
      s->flag = (int *)xcalloc((size_t)s->cap, sizeof(int));
      s->hist = (double *)xcalloc(3 * (size_t)s->cap, sizeof(double));
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
  carry_out(s);
  pbc(s);
  build(s);
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

/* pair loop, pair_gran_base.h:255-430 (newton off: j always local here) */
static void pair_compute(orc_sys *s, int shearupdate) {
  for (int i = 0; i < s->n; i++)
    for (int64_t e = s->first[i]; e < s->first[i + 1]; e++) {
      const int j = s->nbr[e];
      double del[3], dFi[3], dTi[3], dTj[3];
      min_image_del(&s->p, &s->x[3 * i], &s->x[3 * j], del);
      if (orc_contact_pair(&s->p.model, &s->m, s->type[i], s->type[j], del, s->r[i], s->r[j],
                           &s->v[3 * i], &s->v[3 * j], &s->w[3 * i], &s->w[3 * j], s->rm[i],
                           s->rm[j], s->p.dt, shearupdate, &s->hist[3 * e], dFi, dTi, dTj)) {
        /* contact flag set by Hertz / tangential (normal_model_hertz.h:204,
         * tangential_model_history.h:137) */
        s->flag[e] = 1;
        for (int d = 0; d < 3; d++) {
          s->f[3 * i + d] += dFi[d];
          s->t[3 * i + d] += dTi[d];
          s->f[3 * j + d] += -dFi[d];
          s->t[3 * j + d] += dTj[d];
        }
      }
      /* with contact_distance_factor 1 the surfacesClose branch
       * (pair_gran_base.h:414-417) is unreachable: the pair keeps its history */
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
        if (!(deltan <= 0)) continue; /* surfacesClose: no force, history kept */
        const double sdelta[3] = {-delta[0], -delta[1], -delta[2]};
        double dF[3], dT[3];
        orc_contact_wall(&s->p.plane_model[k], &s->m, s->type[i], s->p.plane_type[k], -deltan, sdelta,
                         s->r[i], &s->v[3 * i], &s->w[3 * i], zero, s->rm[i], s->p.dt,
                         shearupdate, h, dF, dT);
        for (int d = 0; d < 3; d++) {
          s->f[3 * i + d] += dF[d];
          s->t[3 * i + d] += dT[d];
        }
      } else {
        h[0] = h[1] = h[2] = 0.0;
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

static void forces(orc_sys *s, int shearupdate) {
  memset(s->f, 0, 3 * sizeof(double) * (size_t)s->n);
  memset(s->t, 0, 3 * sizeof(double) * (size_t)s->n);
  pair_compute(s, shearupdate);
  if (s->p.gravity_first) {
    gravity(s);
    wall_compute(s, shearupdate);
  } else {
    wall_compute(s, shearupdate);
    gravity(s);
  }
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
  forces(s, 0);
}

/* Verlet::run, verlet.cpp:264-392 */
void orc_run(orc_sys *s, int nsteps) {
  for (int k = 0; k < nsteps; k++) {
    s->ntimestep++;
    initial_integrate(s);
    /* Neighbor::decide with delay 0, every 1: check every step */
    if (check_distance(s)) rebuild(s);
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
      flags[e] = s->flag[e];
      for (int d = 0; d < 3; d++) hist[3 * e + d] = sg * s->hist[3 * e + d];
    }
  return np;
}

void orc_get_wall_hist(const orc_sys *s, double *hist) {
  memcpy(hist, s->whist, sizeof(double) * 3 * (size_t)s->n * (size_t)(s->p.nplanes > 0 ? s->p.nplanes : 1));
}

int64_t orc_nrebuilds(const orc_sys *s) { return s->nrebuilds; }
int64_t orc_ntimestep(const orc_sys *s) { return s->ntimestep; }

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
