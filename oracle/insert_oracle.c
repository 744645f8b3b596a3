/*
 * CPU oracle for fix insert/pack (test infrastructure only; see
 * insert_oracle.h). Each function cites the reference file:line it restates.
 */
#include "insert_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define ORI_PI 3.14159265358979323846 /* M_PI */
#define SEED_OFFSET 12                /* fix_insert_pack.cpp:66 */

/* ------------------------------------------------------------------------
 * Park-Miller minimal standard generator, random_park.cpp:54-116
 * ---------------------------------------------------------------------- */
void ori_park_init(ori_park *r, int seed) {
  r->seed = seed;
  r->save = 0;
  r->second = 0.0;
}

double ori_park_uniform(ori_park *r) {
  const int k = r->seed / 127773;
  r->seed = 16807 * (r->seed - k * 127773) - 2836 * k;
  if (r->seed < 0) r->seed += 2147483647;
  return (1.0 / 2147483647) * r->seed;
}

double ori_park_gaussian(ori_park *r) {
  double first, v1, v2, rsq, fac;
  if (!r->save) {
    for (;;) {
      v1 = 2.0 * ori_park_uniform(r) - 1.0;
      v2 = 2.0 * ori_park_uniform(r) - 1.0;
      rsq = v1 * v1 + v2 * v2;
      if (rsq < 1.0 && rsq != 0.0) break;
    }
    fac = sqrt(-2.0 * log(rsq) / rsq);
    r->second = v1 * fac;
    first = v2 * fac;
    r->save = 1;
  } else {
    first = r->second;
    r->save = 0;
  }
  return first;
}

/* ------------------------------------------------------------------------
 * state
 * ---------------------------------------------------------------------- */
struct ori_sys {
  ori_params p;
  ori_park rnd;          /* fix insert's own stream (fix_insert.cpp:100) */
  ori_park region_rnd;   /* the insertion region's stream (region.cpp:446-459, 499-503) */
  ori_park dist_rnd;     /* the distribution's stream (fix_particledistribution_discrete.cpp:83) */
  ori_park tmpl_rnd[ORI_MAXTMPL]; /* random_insertion of each template (fix_template_sphere.cpp:96) */
  double tmpl_vol[ORI_MAXTMPL], tmpl_mass[ORI_MAXTMPL];
  double weight[ORI_MAXTMPL];  /* number-based distribution weights */
  int order[ORI_MAXTMPL];      /* distorder: largest expected volume first */
  double volexpect, massexpect;
  double maxrad;
  double region_volume, region_volume_local;
  /* RegionNeighborList contents: position and radius */
  double *near;
  int nnear, near_cap;
};

/* region block: Region::match (region.cpp:181-189) with RegBlock::inside
 * (region_block.cpp:143-150), interior side */
static int blk_match(const ori_sys *s, const double *x) {
  const double *lo = s->p.region_lo, *hi = s->p.region_hi;
  return x[0] >= lo[0] && x[0] <= hi[0] && x[1] >= lo[1] && x[1] <= hi[1] && x[2] >= lo[2] &&
         x[2] <= hi[2];
}

/* RegBlock::surface_interior (region_block.cpp:159-218): faces within cutoff */
static int blk_surface_interior(const ori_sys *s, const double *x, double cutoff) {
  const double *lo = s->p.region_lo, *hi = s->p.region_hi;
  if (x[0] < lo[0] || x[0] > hi[0] || x[1] < lo[1] || x[1] > hi[1] || x[2] < lo[2] || x[2] > hi[2])
    return 0;
  int n = 0;
  for (int d = 0; d < 3; d++) {
    if (x[d] - lo[d] < cutoff) n++;
    if (hi[d] - x[d] < cutoff) n++;
  }
  return n;
}

/* RegBlock::surface_exterior (region_block.cpp:226-259) with Region::add_contact
 * (region.cpp:248-257) */
static int blk_surface_exterior(const ori_sys *s, const double *x, double cutoff) {
  const double *lo = s->p.region_lo, *hi = s->p.region_hi;
  for (int d = 0; d < 3; d++)
    if (x[d] <= lo[d] - cutoff || x[d] >= hi[d] + cutoff) return 0;
  if (x[0] > lo[0] && x[0] < hi[0] && x[1] > lo[1] && x[1] < hi[1] && x[2] > lo[2] && x[2] < hi[2])
    return 0;
  double p[3];
  for (int d = 0; d < 3; d++) p[d] = x[d] < lo[d] ? lo[d] : (x[d] > hi[d] ? hi[d] : x[d]);
  const double dx = x[0] - p[0], dy = x[1] - p[1], dz = x[2] - p[2];
  const double r = sqrt(dx * dx + dy * dy + dz * dz);
  return r < cutoff;
}

/* Region::match_cut (region.cpp:598-606) and match_expandby_cut (:608-619) */
static int blk_match_cut(const ori_sys *s, const double *x, double cut) {
  return blk_surface_interior(s, x, cut);
}
static int blk_match_expandby_cut(const ori_sys *s, const double *x, double cut) {
  return blk_match(s, x) || blk_surface_exterior(s, x, cut);
}

/* Region::rand_bounds (region.cpp:505-527), subdomain = the whole box */
static int rand_bounds(const ori_sys *s, const double *boxlo, const double *boxhi, double *lo,
                       double *hi) {
  for (int d = 0; d < 3; d++) {
    lo[d] = s->p.region_lo[d] > boxlo[d] ? s->p.region_lo[d] : boxlo[d];
    hi[d] = s->p.region_hi[d] < boxhi[d] ? s->p.region_hi[d] : boxhi[d];
  }
  return !(lo[0] >= hi[0] || lo[1] >= hi[1] || lo[2] >= hi[2]);
}

/* Region::generate_random (region.cpp:531-544) */
static void gen_random(ori_sys *s, const double *lo, const double *diff, double *pos) {
  do {
    pos[0] = lo[0] + ori_park_uniform(&s->region_rnd) * diff[0];
    pos[1] = lo[1] + ori_park_uniform(&s->region_rnd) * diff[1];
    pos[2] = lo[2] + ori_park_uniform(&s->region_rnd) * diff[2];
  } while (!blk_match(s, pos));
}

/* Region::generate_random_shrinkby_cut (region.cpp:550-571) */
static void gen_random_shrinkby_cut(ori_sys *s, const double *lo, const double *diff, double cut,
                                    double *pos) {
  do {
    pos[0] = lo[0] + ori_park_uniform(&s->region_rnd) * diff[0];
    pos[1] = lo[1] + ori_park_uniform(&s->region_rnd) * diff[1];
    pos[2] = lo[2] + ori_park_uniform(&s->region_rnd) * diff[2];
  } while (!blk_match(s, pos) || blk_match_cut(s, pos, cut));
}

/* vectorMin3D (vector_liggghts.h:176-185) */
static double min3(const double *v) {
  if (v[0] < v[1] && v[0] < v[2]) return v[0];
  if (v[1] < v[2]) return v[1];
  return v[2];
}

/* Domain::dist_subbox_borders (domain_I.h:126-143), one process */
static double dist_box_borders(const double *boxlo, const double *boxhi, const double *pos) {
  double dl[3], dh[3];
  for (int d = 0; d < 3; d++) {
    dl[d] = boxlo[d] - pos[d];
    dh[d] = boxhi[d] - pos[d];
    if (dl[d] < 0) dl[d] = -dl[d];
    if (dh[d] < 0) dh[d] = -dh[d];
  }
  const double dlo = min3(dl), dhi = min3(dh);
  return dlo < dhi ? dlo : dhi;
}

/* Region::volume_mc (region.cpp:634-701), one process; the domain test is
 * Domain::is_in_domain (domain_I.h:51-62) and is_in_subdomain is the same box */
static int volume_mc(ori_sys *s, int n_test, int cutflag, double cut, const double *blo,
                     const double *bhi) {
  const double *lo = s->p.region_lo, *hi = s->p.region_hi;
  int n_in_local = 0, n_in_global = 0;
  for (int i = 0; i < n_test; i++) {
    double pos[3];
    pos[0] = lo[0] + ori_park_uniform(&s->region_rnd) * (hi[0] - lo[0]);
    pos[1] = lo[1] + ori_park_uniform(&s->region_rnd) * (hi[1] - lo[1]);
    pos[2] = lo[2] + ori_park_uniform(&s->region_rnd) * (hi[2] - lo[2]);
    if (!(pos[0] >= blo[0] && pos[0] <= bhi[0] && pos[1] >= blo[1] && pos[1] <= bhi[1] &&
          pos[2] >= blo[2] && pos[2] <= bhi[2]))
      continue;
    if (blk_match(s, pos)) {
      n_in_global++;
      /* is_in_subdomain: [lo, hi) padded at the box faces, which every point
       * already inside the box satisfies */
      if (!cutflag || !blk_match_cut(s, pos, cut)) n_in_local++;
    }
  }
  if (n_in_global == 0) return 0;
  const double vol_bbox = (hi[0] - lo[0]) * (hi[1] - lo[1]) * (hi[2] - lo[2]);
  const double vol_global = (double)n_in_global / (double)(n_test * 1) * vol_bbox;
  double vol_local = (double)n_in_local / (double)n_test * vol_bbox;
  const double vol_local_all = vol_local;
  if (vol_local_all < 1.e-10) return 0; /* volume_limit_ default, region.cpp:93 */
  vol_local *= (vol_global / vol_local_all);
  s->region_volume = vol_global;
  s->region_volume_local = vol_local;
  return 1;
}

ori_sys *ori_create(const ori_params *p) {
  if (p->ntemplates < 1 || p->ntemplates > ORI_MAXTMPL) return NULL;
  ori_sys *s = (ori_sys *)calloc(1, sizeof(ori_sys));
  if (!s) return NULL;
  s->p = *p;
  const int nt = p->ntemplates;
  /* FixTemplateSphere (fix_template_sphere.cpp:96-245): constant radius and
   * density; volume_expect = r^3 * 4 pi / 3, mass_expect = rho * volume */
  for (int i = 0; i < nt; i++) {
    ori_park_init(&s->tmpl_rnd[i], p->tmpl_seed[i]);
    const double r = p->tmpl_radius[i];
    s->tmpl_vol[i] = r * r * r * 4. * ORI_PI / 3.;
    s->tmpl_mass[i] = p->tmpl_density[i] * s->tmpl_vol[i];
  }
  /* FixParticledistributionDiscrete (fix_particledistribution_discrete.cpp:69-244) */
  ori_park_init(&s->dist_rnd, p->dist_seed);
  double wsum = 0;
  for (int i = 0; i < nt; i++) wsum += p->dist_weight[i];
  for (int i = 0; i < nt; i++) s->weight[i] = p->dist_weight[i] / wsum;
  if (!p->dist_numberbased) {
    for (int i = 0; i < nt; i++) s->weight[i] = s->weight[i] / s->tmpl_mass[i];
    wsum = 0;
    for (int i = 0; i < nt; i++) wsum += s->weight[i];
    for (int i = 0; i < nt; i++) s->weight[i] /= wsum;
  }
  s->volexpect = s->massexpect = 0.;
  for (int i = 0; i < nt; i++) {
    s->volexpect += s->tmpl_vol[i] * s->weight[i];
    s->massexpect += s->tmpl_mass[i] * s->weight[i];
  }
  for (int i = 0; i < nt; i++) s->order[i] = i;
  int swapped, n = nt;
  do {
    swapped = 0;
    for (int i = 0; i < nt - 1; i++)
      if (s->tmpl_vol[s->order[i]] < s->tmpl_vol[s->order[i + 1]]) {
        const int t = s->order[i];
        s->order[i] = s->order[i + 1];
        s->order[i + 1] = t;
        swapped = 1;
      }
    n--;
  } while (swapped && n > 0);
  s->maxrad = 0.;
  for (int i = 0; i < nt; i++)
    if (p->tmpl_radius[i] > s->maxrad) s->maxrad = p->tmpl_radius[i];
  /* FixInsert (fix_insert.cpp:98-101) and FixInsertPack::calc_insertion_properties
   * (fix_insert_pack.cpp:184-222): exactly one target */
  ori_park_init(&s->rnd, p->seed);
  int ndef = (p->volumefraction_region > 0.) + (p->particles_in_region > 0) + (p->mass_in_region > 0.);
  if (ndef != 1) {
    free(s);
    return NULL;
  }
  ori_park_init(&s->region_rnd, p->seed + SEED_OFFSET);
  if (!volume_mc(s, p->ntry_mc, p->all_in, s->maxrad, p->boxlo, p->boxhi) || s->region_volume <= 0. ||
      s->region_volume_local < 0. ||
      (s->region_volume_local - s->region_volume) / s->region_volume > 1e-3) {
    free(s);
    return NULL;
  }
  return s;
}

void ori_destroy(ori_sys *s) {
  if (!s) return;
  free(s->near);
  free(s);
}

double ori_region_volume(const ori_sys *s) { return s->region_volume; }

void ori_distribution(const ori_sys *s, double *volexpect, double *massexpect, double *weights,
                      int *order) {
  if (volexpect) *volexpect = s->volexpect;
  if (massexpect) *massexpect = s->massexpect;
  for (int i = 0; i < s->p.ntemplates; i++) {
    if (weights) weights[i] = s->weight[i];
    if (order) order[i] = s->order[i];
  }
}

void ori_rng_states(const ori_sys *s, int *st) {
  st[0] = s->rnd.seed;
  st[1] = s->region_rnd.seed;
  st[2] = s->dist_rnd.seed;
  st[3] = s->tmpl_rnd[0].seed;
}

static void near_add(ori_sys *s, const double *x, double r) {
  if (s->nnear == s->near_cap) {
    s->near_cap = s->near_cap ? 2 * s->near_cap : 1024;
    s->near = (double *)realloc(s->near, 4 * sizeof(double) * (size_t)s->near_cap);
    if (!s->near) abort();
  }
  double *q = s->near + 4 * (size_t)s->nnear++;
  q[0] = x[0];
  q[1] = x[1];
  q[2] = x[2];
  q[3] = r;
}

/* RegionNeighborList::hasOverlap (region_neighbor_list_I.h:72-98), all listed particles */
static int has_overlap(const ori_sys *s, const double *x, double radius) {
  for (int k = 0; k < s->nnear; k++) {
    const double *q = s->near + 4 * (size_t)k;
    const double d0 = x[0] - q[0], d1 = x[1] - q[1], d2 = x[2] - q[2];
    const double rsq = d0 * d0 + d1 * d1 + d2 * d2;
    const double radsum = radius + q[3];
    if (rsq <= radsum * radsum) return 1;
  }
  return 0;
}

int ori_insert_step(ori_sys *s, int n, const double *x, const double *radius, const double *rmass,
                    const double *boxlo, const double *boxhi, int cap, double *xo, double *vo,
                    double *wo, double *ro, double *rho, double *mo, int *typeo, int *requested) {
  const ori_params *p = &s->p;
  const int nt = p->ntemplates;
  if (requested) *requested = 0;
  /* FixInsertPack::calc_ninsert_this (fix_insert_pack.cpp:236-330) */
  int np_region = 0;
  double vol_region = 0., mass_region = 0.;
  const double _4Pi3 = 4. * ORI_PI / 3.;
  for (int i = 0; i < n; i++)
    if (blk_match(s, x + 3 * (size_t)i)) {
      np_region++;
      vol_region += _4Pi3 * radius[i] * radius[i] * radius[i];
      mass_region += rmass[i];
    }
  int ninsert_this;
  double ratio;
  if (p->volumefraction_region > 0.) {
    ninsert_this = (int)((p->volumefraction_region * s->region_volume - vol_region) / s->volexpect +
                         ori_park_uniform(&s->rnd));
    ratio = vol_region / (p->volumefraction_region * s->region_volume);
  } else if (p->particles_in_region > 0) {
    ninsert_this = p->particles_in_region - np_region;
    ratio = (double)np_region / (double)p->particles_in_region;
  } else {
    ninsert_this = (int)((p->mass_in_region - mass_region) / s->massexpect + ori_park_uniform(&s->rnd));
    ratio = mass_region / p->mass_in_region;
  }
  if (ninsert_this < -200000) return -1;
  if (ninsert_this < 0) ninsert_this = 0;
  if (ratio < 0.) ratio = 0.;
  if (ratio > 1.) ratio = 1.;

  /* FixInsert::distribute_ninsert_this (fix_insert.cpp:829-913), one process:
   * the exact split gives everything to it and draws nothing. It first asks
   * for FixInsertPack::insertion_fraction (fix_insert_pack.cpp:356-362), which
   * measures the region volume again when the box can change. */
  if (p->box_change && !volume_mc(s, p->ntry_mc, p->all_in, s->maxrad, boxlo, boxhi)) return -1;
  int nlocal = ninsert_this;
  if (!p->exact_number) {
    const double fraction_local = s->region_volume_local / s->region_volume;
    nlocal = (int)(fraction_local * (double)ninsert_this + ori_park_uniform(&s->rnd));
  }

  /* FixParticledistributionDiscrete::randomize_list
   * (fix_particledistribution_discrete.cpp:395-469) */
  int parttogen[ORI_MAXTMPL];
  if (!p->exact_number) {
    for (int i = 0; i < nt; i++)
      parttogen[i] = (int)((double)nlocal * s->weight[i] + ori_park_uniform(&s->dist_rnd));
  } else {
    double rem[ORI_MAXTMPL];
    int trunc = 0;
    for (int i = 0; i < nt; i++) {
      parttogen[i] = (int)((double)nlocal * s->weight[i]);
      trunc += parttogen[i];
      rem[i] = (double)nlocal * s->weight[i] - (double)parttogen[i];
    }
    const int gap = nlocal - trunc;
    for (int g = 0; g < gap; g++) {
      const double r = ori_park_uniform(&s->dist_rnd) * (double)gap;
      int j = 0;
      double rsum = rem[0];
      while (rsum < r && j < nt - 1) {
        j++;
        rsum += rem[j];
      }
      parttogen[j]++;
    }
  }
  /* the insertion list: templates in distorder, each template's particles in a row */
  int nlist = 0;
  for (int i = 0; i < nt; i++) nlist += parttogen[i];
  if (requested) *requested = nlist;
  if (nlist == 0) return 0;
  int *list_tmpl = (int *)malloc(sizeof(int) * (size_t)nlist);
  if (!list_tmpl) abort();
  {
    int k = 0;
    for (int i = 0; i < nt; i++) {
      const int d = s->order[i];
      for (int j = 0; j < parttogen[d]; j++) list_tmpl[k++] = d;
    }
  }

  /* FixInsert::load_xnear (fix_insert.cpp:916-948) with FixInsertPack::is_nearby
   * (fix_insert_pack.cpp:345-360) and getBoundingBox (:364-374) */
  s->nnear = 0;
  double bblo[3], bbhi[3];
  for (int d = 0; d < 3; d++) {
    bblo[d] = p->region_lo[d] < boxlo[d] ? boxlo[d] : p->region_lo[d];
    bbhi[d] = p->region_hi[d] > boxhi[d] ? boxhi[d] : p->region_hi[d];
    bblo[d] -= 2 * s->maxrad;
    bbhi[d] += 2 * s->maxrad;
  }
  if (p->overlapcheck) {
    for (int i = 0; i < n; i++) {
      const double *xi = x + 3 * (size_t)i;
      const double cut = p->all_in ? s->maxrad : radius[i] + s->maxrad;
      if (!blk_match_expandby_cut(s, xi, cut)) continue;
      if (xi[0] >= bblo[0] && xi[0] <= bbhi[0] && xi[1] >= bblo[1] && xi[1] <= bbhi[1] &&
          xi[2] >= bblo[2] && xi[2] <= bbhi[2])
        near_add(s, xi, radius[i]);
    }
  }

  /* FixInsertPack::x_v_omega (fix_insert_pack.cpp:390-500) */
  double lo[3], hi[3], diff[3];
  if (!rand_bounds(s, boxlo, boxhi, lo, hi)) {
    free(list_tmpl);
    return -1;
  }
  for (int d = 0; d < 3; d++) diff[d] = hi[d] - lo[d];
  int maxtry;
  if (ratio >= 1.) maxtry = nlist * p->maxattempt;
  else maxtry = (int)((double)(nlist * p->maxattempt) / (1. - ratio));
  int ntry = 0, nins_total = 0;
  for (;;) {
    if (p->overlapcheck && !(ntry < maxtry && nins_total < nlist)) break;
    if (!p->overlapcheck && nins_total >= nlist) break;
    const int t = list_tmpl[nins_total];
    const double rbound = p->tmpl_radius[t];
    double pos[3], v[3];
    int ok = 0;
    if (!p->overlapcheck) {
      if (p->all_in) gen_random_shrinkby_cut(s, lo, diff, rbound, pos);
      else gen_random(s, lo, diff, pos);
      ok = 1;
    } else {
      while (!ok && ntry < maxtry) {
        do {
          if (p->all_in) gen_random_shrinkby_cut(s, lo, diff, rbound, pos);
          else gen_random(s, lo, diff, pos);
          ntry++;
        } while (p->check_dist_from_subdomain_border && ntry < maxtry &&
                 dist_box_borders(boxlo, boxhi, pos) < rbound);
        if (ntry == maxtry) break;
        /* FixInsert::generate_random_velocity (fix_insert.cpp:956-972) */
        for (int d = 0; d < 3; d++) v[d] = p->v_insert[d];
        if (p->vel_style == 1)
          for (int d = 0; d < 3; d++)
            v[d] = p->v_insert[d] + p->v_fluct[d] * 2.0 * (ori_park_uniform(&s->rnd) - 0.50);
        else if (p->vel_style == 2)
          for (int d = 0; d < 3; d++) v[d] = p->v_insert[d] + p->v_fluct[d] * ori_park_gaussian(&s->rnd);
        /* ParticleToInsert::check_near_set_x_v_omega (particleToInsert.cpp:247-265) */
        if (has_overlap(s, pos, rbound)) continue;
        near_add(s, pos, rbound);
        ok = 1;
      }
      if (!ok) continue; /* ntry reached maxtry: the outer test ends the loop */
    }
    if (p->overlapcheck == 0) {
      for (int d = 0; d < 3; d++) v[d] = p->v_insert[d];
      if (p->vel_style == 1)
        for (int d = 0; d < 3; d++)
          v[d] = p->v_insert[d] + p->v_fluct[d] * 2.0 * (ori_park_uniform(&s->rnd) - 0.50);
      else if (p->vel_style == 2)
        for (int d = 0; d < 3; d++) v[d] = p->v_insert[d] + p->v_fluct[d] * ori_park_gaussian(&s->rnd);
    }
    /* ParticleToInsert::insert (particleToInsert.cpp:99-147) with the template's
     * values (fix_template_sphere.cpp:296-321) */
    const int k = nins_total++;
    if (k < cap) {
      const double r = p->tmpl_radius[t];
      for (int d = 0; d < 3; d++) {
        xo[3 * (size_t)k + d] = pos[d];
        vo[3 * (size_t)k + d] = v[d];
        wo[3 * (size_t)k + d] = p->omega[d];
      }
      ro[k] = r;
      rho[k] = p->tmpl_density[t];
      mo[k] = p->tmpl_density[t] * (r * r * r * 4. * ORI_PI / 3.);
      typeo[k] = p->tmpl_type[t];
    }
  }
  free(list_tmpl);
  return nins_total;
}
