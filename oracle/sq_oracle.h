/* CPU oracle, test infrastructure only: superquadric contact-point solve
 * (see sq_oracle.c for the reference functions it restates). */
#ifndef SQ_ORACLE_H
#define SQ_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* One superquadric: centre, orientation quaternion (w, i, j, k), half-axes
 * a, b, c and blockiness n1, n2 (atom->shape / atom->blockiness). */
typedef struct {
  double center[3];
  double quat[4];
  double shape[3];
  double blockiness[2];
} orc_sq_pose;

/* Shape function value, gradient and Hessian (may be NULL) at a global point. */
void orc_sq_props(const orc_sq_pose *p, const double *x, double *f, double *grad, double *hess);

/* Contact point of A and B. init == NULL: a fresh contact (the homotopy from
 * spheres); otherwise one solve from init. Returns the fail flag; *iters is
 * the total number of Newton iterations. */
int orc_sq_contact(const orc_sq_pose *a, const orc_sq_pose *b, const double *init, double *point,
                   double *fi, double *fj, int *iters);

/* n independent contacts; init may be NULL (all fresh) or n*3 points. */
void orc_sq_contact_batch(int64_t n, const orc_sq_pose *a, const orc_sq_pose *b, const double *init,
                          double *point, double *fi, double *fj, int *iters, int *fail);

/* SurfaceModel<SURFACE_SUPERQUADRIC>::checkSurfaceIntersect for one pair:
 * hist[7] = state (0 far, 1 close, 2 intersecting), first OBB inequality,
 * contact point[3], alpha_i, alpha_j (in/out); vel = v_i, v_j, omega_i,
 * omega_j (12) and mass = m_i, m_j may be NULL (a still pair). Returns 1 when
 * the surfaces intersect, with the normal en (from i to j), the overlap
 * deltan, the surface points on i and j, delta = cpb - cpa. */
int orc_sq_check(const orc_sq_pose *a, const orc_sq_pose *b, double *hist, const double *vel,
                 const double *mass, double dt, double *en, double *deltan, double *cpa, double *cpb,
                 double *delta, int *iters);
void orc_sq_check_batch(int64_t n, const orc_sq_pose *a, const orc_sq_pose *b, double *hist, const double *vel,
                        const double *mass, double dt, int *contact, double *en, double *deltan, double *cpa,
                        double *cpb, int *iters);
/* Superquadric::plane_intersection: the particle's point of maximal
 * penetration and the wall's point of lowest potential; returns 1 when they
 * intersect. */
int orc_sq_plane(const orc_sq_pose *p, const double *normal_global, const double *x0_global, double *pmax,
                 double *plow);
/* bounding_sphere_radius_superquadric, times scale_koef. */
double orc_sq_bounding_radius(const double *shape, const double *blk);
/* obb_intersect with the first inequality to try (in/out). */
int orc_sq_obb(const orc_sq_pose *a, const orc_sq_pose *b, int *start);
/* Superquadric::surface_line_intersection (20 iterations, no ellipsoid start). */
double orc_sq_line(const orc_sq_pose *p, const double *start, const double *dir, double alpha_prev,
                   double *result);

#ifdef __cplusplus
}
#endif

#endif
