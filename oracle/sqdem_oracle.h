/*
 * CPU oracle for the superquadric DEM step, test infrastructure only: a
 * plain-C FP64 restatement of LIGGGHTS-PUBLIC 3.8.0's per-timestep path for
 * atom_style superquadric with pair_style gran model hertz tangential history
 * surface superquadric, fix wall/gran ... primitive, fix gravity and
 * fix nve/superquadric. Nothing in the product links or calls it.
 *
 * Ghost atoms are replaced by periodic images of the partner (x_j + k L with
 * k in {-1, 0, 1}, chosen by the minimum image); each pair is stored once,
 * i < j by index.
 */
#ifndef SQDEM_ORACLE_H
#define SQDEM_ORACLE_H

#include <stdint.h>

#include "dem_oracle.h"
#include "sq_oracle.h"

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_SQS_HIST 10 /* state, first OBB inequality, contact point[3], alpha_i, alpha_j, shear[3] */

typedef struct {
  double dt;
  double skin;
  double boxlo[3], boxhi[3];
  int periodic[3];
  int tangential_damping; /* default on */
  int limit_force;        /* default off */
  int has_gravity;
  double grav_acc[3];
  int nplanes;
  int plane_axis[ORC_MAXPLANE];
  double plane_pos[ORC_MAXPLANE];
  int plane_type[ORC_MAXPLANE];
  int gravity_first;
} orc_sqs_params;

typedef struct orc_sqs orc_sqs;

/* volume_superquadric / inertia_superquadric (math_extra_liggghts_superquadric.cpp:
 * 242-268): volume, and principal moments of inertia for the given density */
void orc_sqs_mass_props(const double *shape, const double *blk, double density, double *volume,
                        double *inertia);

/* One superquadric pair contact (checkSurfaceIntersect, then the Hertz /
 * tangential history chain with contact-point lever arms). hist
 * (ORC_SQS_HIST) is read and written; dFi / dTi / dTj are the force on i and
 * the torques on i and j (force on j is -dFi). Returns 1 on contact. */
int orc_sqs_contact_pair(const orc_materials *mat, int tangential_damping, int limit_force, double dt,
                         int itype, int jtype, const orc_sq_pose *a, const orc_sq_pose *b, const double *vi,
                         const double *vj, const double *wi, const double *wj, double mi, double mj, double rvi,
                         double rvj, int shearupdate, double *hist, double *dFi, double *dTi, double *dTj);

/* n particles: x, v, omega (n*3), quat (n*4, w first), shape (n*3),
 * blockiness (n*2), density (n), type (n, 1-based). */
orc_sqs *orc_sqs_create(const orc_sqs_params *p, const orc_materials *m, int n, const double *x,
                        const double *v, const double *omega, const double *quat, const double *shape,
                        const double *blk, const double *density, const int *type);
void orc_sqs_destroy(orc_sqs *s);
/* Verlet::setup: wrap, build, forces without history update */
void orc_sqs_setup(orc_sqs *s);
void orc_sqs_run(orc_sqs *s, int nsteps);
void orc_sqs_get_state(const orc_sqs *s, double *x, double *v, double *omega, double *angmom, double *quat,
                       double *f, double *torque);
/* mass, volume-equivalent radius, bounding radius (n each), inertia (n*3) */
void orc_sqs_get_props(const orc_sqs *s, double *mass, double *rvol, double *rbound, double *inertia);
/* pairs of the list (i < j), touch flags and histories (ORC_SQS_HIST each) */
int64_t orc_sqs_get_pairs(const orc_sqs *s, int64_t cap, int *ij, int *flags, double *hist);
void orc_sqs_get_wall_hist(const orc_sqs *s, double *hist /* n*nplanes*3 */);
/* touching pairs, touching particle-plane contacts, Newton iterations of the
 * last force evaluation, rebuilds so far, steps so far */
void orc_sqs_stats(const orc_sqs *s, int64_t *npair_touch, int64_t *nwall_touch, int64_t *newton_iters,
                   int64_t *nrebuild, int64_t *ntimestep);

#ifdef __cplusplus
}
#endif

#endif
