/*
 * CPU oracle for particle insertion: a plain-C restatement of LIGGGHTS-PUBLIC
 * 3.8.0's fix insert/pack with fix particletemplate/sphere and
 * fix particledistribution/discrete on a block region, one process.
 *
 * TEST INFRASTRUCTURE ONLY. Nothing in the shipped library links, loads or
 * calls this code; only tests/ use it, as the checker.
 *
 * The overlap check is a brute-force scan of every particle the reference
 * puts in its RegionNeighborList (existing particles near the region and the
 * ones inserted so far). The reference bins them (bins of 4 * maxrad, a
 * 27-bin stencil, region_neighbor_list_I.h:403-560) and scans the bins around
 * the candidate (:72-98); the two agree whenever no listed particle is larger
 * than the largest template radius, which holds for every case the tests run.
 */
#ifndef INSERT_ORACLE_H
#define INSERT_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORI_MAXTMPL 8

/* random_park.cpp:54-116 */
typedef struct {
  int seed;
  int save;
  double second;
} ori_park;

void ori_park_init(ori_park *r, int seed);
double ori_park_uniform(ori_park *r);
double ori_park_gaussian(ori_park *r);

typedef struct {
  /* fix <id> all particletemplate/sphere <seed> atom_type <t>
   *     density constant <rho> radius constant <r> (fix_template_sphere.cpp:72-245) */
  int ntemplates;
  int tmpl_seed[ORI_MAXTMPL];
  int tmpl_type[ORI_MAXTMPL];
  double tmpl_radius[ORI_MAXTMPL];
  double tmpl_density[ORI_MAXTMPL];
  /* fix <id> all particledistribution/discrete[/numberbased] <seed> <n> id w ...
   * (fix_particledistribution_discrete.cpp:69-244) */
  int dist_seed;
  int dist_numberbased;
  double dist_weight[ORI_MAXTMPL];
  /* fix <id> <grp> insert/pack seed <s> distributiontemplate <id> ... region <rid>
   * (fix_insert.cpp:78-340, fix_insert_pack.cpp:74-160) */
  int seed;
  int maxattempt;          /* default 50 */
  int overlapcheck;        /* default yes */
  int all_in;              /* default no */
  int exact_number;        /* random_distribute exact (1, default) | uncorrelated (0) */
  int vel_style;           /* 0 constant, 1 uniform, 2 gaussian */
  double v_insert[3], v_fluct[3], omega[3];
  int check_dist_from_subdomain_border; /* default yes */
  int ntry_mc;             /* default 100000 */
  double volumefraction_region;
  int particles_in_region;
  double mass_in_region;
  /* region <rid> block xlo xhi ylo yhi zlo zhi (region_block.cpp:60-130) */
  double region_lo[3], region_hi[3];
  /* simulation box at the first run's setup (Region::volume_mc's is_in_domain) */
  double boxlo[3], boxhi[3];
  /* Domain::box_change (shrink-wrapped faces, domain.cpp:192-208): the region
   * volume is measured again with the current box at every insertion */
  int box_change;
} ori_params;

typedef struct ori_sys ori_sys;

/* Constructors and FixInsertPack::calc_insertion_properties (the first run's
 * setup): region random generator reset to seed + 12, region volume by Monte
 * Carlo. Returns NULL on an invalid setting (the reference's error->all). */
ori_sys *ori_create(const ori_params *p);
void ori_destroy(ori_sys *s);
double ori_region_volume(const ori_sys *s);
/* expected volume / mass per inserted particle, the number-based weights and
 * the insertion order of the templates (largest expected volume first) */
void ori_distribution(const ori_sys *s, double *volexpect, double *massexpect, double *weights,
                      int *order);

/* One FixInsert::pre_exchange insertion (fix_insert.cpp:652-826) with the
 * current particles (x n*3, radius, rmass) and the current box. Writes up to
 * cap inserted particles in insertion order: position, velocity, angular
 * velocity (3 each), radius, density, mass, atom type. Returns the number
 * inserted; *requested receives the number the fix asked for (after the
 * distribution's rounding). */
int ori_insert_step(ori_sys *s, int n, const double *x, const double *radius, const double *rmass,
                    const double *boxlo, const double *boxhi, int cap, double *xo, double *vo,
                    double *wo, double *ro, double *rho, double *mo, int *typeo, int *requested);

/* state of the four random streams (fix, region, distribution, first template) */
void ori_rng_states(const ori_sys *s, int *states);

#ifdef __cplusplus
}
#endif

#endif
