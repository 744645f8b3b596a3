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

#ifdef __cplusplus
}
#endif

#endif
