/*
 * CPU oracle, test infrastructure only: plain-C FP64 restatement of the
 * superquadric contact-point solve of LIGGGHTS-PUBLIC 3.8.0. Nothing in the
 * product links or calls this file; tests and the benchmark's CPU baseline
 * use it as the checker.
 *
 * Follows:
 *   Superquadric::set / set_shape / set_blockiness / calc_koef
 *       (superquadric.cpp:84-110, 673-703)
 *   Superquadric::shape_function_props_local / _global
 *       (superquadric.cpp:237-326), tensor_quat_rotate (:620-625)
 *   MathExtraLiggghtsNonspherical::pow_abs_int / pow_abs / sign / isInteger /
 *       quat_to_mat / matvecN / GMRES<4,4> (math_extra_liggghts_nonspherical.h:
 *       137-324, 511-534, 578-582)
 *   calc_F (both variants), determinant_4x4, inverseMatrix4x4, the Newton /
 *       GMRES / golden-section solve calc_contact_point
 *       (math_extra_liggghts_superquadric.cpp:619-658, 715-1022)
 *   calc_contact_point_if_no_previous_point_avaialable: the homotopy from
 *       spheres to the actual shapes that gives a fresh contact its first
 *       point (math_extra_liggghts_superquadric.cpp:1390-1430)
 * Each solve also reports how many Newton iterations it took.
 */
#include "sq_oracle.h"

#include <math.h>
#include <string.h>

#define PHI_INV 0.61803398874989479

typedef struct {
  double center[3], quat[4], shape[3], blockiness[2];
  double rot[9], koef, shape_inv[3];
  int int_blockiness;
  double gradient[3], hessian[9];
} sq_t;

/* MathExtraLiggghts::compDouble (math_extra_liggghts.h:536-547) */
static int comp_double(double a, double b, double prec) {
  if (a == b) return 1;
  if (b == 0) return a < prec && a > -prec;
  const double x = a - b;
  return x < prec && x > -prec;
}

/* pow_abs_int (math_extra_liggghts_nonspherical.h:137-193) */
static double pow_abs_int(double a, int b) {
  a = fabs(a);
  switch (b) {
    case 0: return 1.0;
    case 1: return a;
    case 2: return a * a;
    case 3: return a * a * a;
    case 4: { const double a2 = a * a; return a2 * a2; }
    case 5: { const double a2 = a * a; return a2 * a2 * a; }
    case 6: { const double a3 = a * a * a; return a3 * a3; }
    case 7: { const double a2 = a * a, a3 = a2 * a; return a2 * a2 * a3; }
    case 8: { const double a2 = a * a, a4 = a2 * a2; return a4 * a4; }
    case 9: { const double a3 = a * a * a; return a3 * a3 * a3; }
    default: {
      double r = a;
      for (int i = 0; i < b - 1; i++) r *= a;
      return r;
    }
  }
}

/* pow_abs (math_extra_liggghts_nonspherical.h:195-221) */
static double pow_abs(double base, double exponent) {
  base = fabs(base);
  const double tol = 1e-12;
  if (fabs(exponent) < tol || fabs(base - 1.0) < tol) return 1.0;
  if (fabs(base) < tol) return 0.0;
  const double n = floor(fabs(exponent));
  double r = pow_abs_int(base, (int)n);
  const double nd = fabs(exponent) - n;
  if (nd > 0.0) {
    if (fabs(2.0 * nd - 1.0) < 2.0 * tol)
      r *= sqrt(base);
    else
      r *= pow(base, nd);
  }
  return exponent > 0.0 ? r : 1.0 / r;
}

/* sign with a dead band (math_extra_liggghts_nonspherical.h:316-324) */
static double sgn(double v) {
  if (v > 1e-14) return 1.0;
  if (v < -1e-14) return -1.0;
  return 0.0;
}

/* isInteger (math_extra_liggghts_nonspherical.h:578-582) */
static int is_integer(double x) { return comp_double(x, floor(x + 1e-2), 1e-2); }

/* quat_to_mat (math_extra_liggghts_nonspherical.h:511-534), quat = (w, i, j, k) */
static void quat_to_mat(const double *q, double *m) {
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

static void sq_set_blockiness(sq_t *s, double n1, double n2) {
  s->blockiness[0] = n1;
  s->blockiness[1] = n2;
  s->koef = pow_abs(0.5, (n1 > n2 ? n1 : n2) - 2.0); /* calc_koef, superquadric.cpp:701-703 */
  const int ellipsoid = comp_double(n1, 2.0, 1e-2) && comp_double(n2, 2.0, 1e-2);
  s->int_blockiness = ellipsoid || (is_integer(n1) && is_integer(n2));
}

static void sq_set_shape(sq_t *s, double a, double b, double c) {
  s->shape[0] = a;
  s->shape[1] = b;
  s->shape[2] = c;
  s->shape_inv[0] = 1.0 / a;
  s->shape_inv[1] = 1.0 / b;
  s->shape_inv[2] = 1.0 / c;
}

static void sq_init(sq_t *s, const orc_sq_pose *p) {
  memcpy(s->center, p->center, sizeof s->center);
  memcpy(s->quat, p->quat, sizeof s->quat);
  quat_to_mat(s->quat, s->rot);
  sq_set_shape(s, p->shape[0], p->shape[1], p->shape[2]);
  sq_set_blockiness(s, p->blockiness[0], p->blockiness[1]);
  memset(s->gradient, 0, sizeof s->gradient);
  memset(s->hessian, 0, sizeof s->hessian);
}

/* shape_function_props_local (superquadric.cpp:237-310) */
static void props_local(const sq_t *s, const double *x, double *f, double *grad, double *hess) {
  const double n1 = s->blockiness[0], n2 = s->blockiness[1];
  const double a = s->shape_inv[0], b = s->shape_inv[1], c = s->shape_inv[2];
  const double xa = x[0] * a, yb = x[1] * b, zc = x[2] * c;
  double xan22, ybn22, zcn12;
  if (s->int_blockiness) {
    const int n1i = (int)floor(n1 + 1e-2), n2i = (int)floor(n2 + 1e-2);
    xan22 = pow_abs_int(xa, n2i - 2);
    ybn22 = pow_abs_int(yb, n2i - 2);
    zcn12 = pow_abs_int(zc, n1i - 2);
  } else {
    xan22 = pow_abs(xa, n2 - 2.0);
    ybn22 = pow_abs(yb, n2 - 2.0);
    zcn12 = pow_abs(zc, n1 - 2.0);
  }
  const double xan21 = xan22 * fabs(xa), ybn21 = ybn22 * fabs(yb), zcn11 = zcn12 * fabs(zc);
  const double xan2 = xan21 * fabs(xa), ybn2 = ybn21 * fabs(yb), zcn1 = zcn11 * fabs(zc);
  const double k = s->koef;
  if (comp_double(n1, n2, 1e-2)) {
    *f = k * (xan2 + ybn2 + zcn1 - 1.0);
    grad[0] = k * a * n1 * sgn(xa) * xan21;
    grad[1] = k * b * n1 * sgn(yb) * ybn21;
    grad[2] = k * c * n1 * sgn(zc) * zcn11;
    if (hess) {
      hess[0] = (k * a) * (n1 * (n2 - 1.0) * xan22) * a;
      hess[4] = (k * b) * (n1 * (n2 - 1.0) * ybn22) * b;
      hess[8] = (k * c) * (n1 * (n1 - 1.0) * zcn12) * c;
      hess[1] = hess[3] = hess[2] = hess[6] = hess[5] = hess[7] = 0.0;
    }
  } else {
    *f = k * (pow_abs(xan2 + ybn2, n1 / n2) + zcn1 - 1.0);
    const double xy = xan21 * fabs(xa) + ybn21 * fabs(yb);
    double xyp2;
    if (is_integer(n1 / n2))
      xyp2 = pow_abs_int(xy, (int)((double)(int)floor(n1 / n2 + 1e-2) - 2.0));
    else
      xyp2 = pow_abs(xy, n1 / n2 - 2.0);
    const double xyp1 = xyp2 * xy;
    grad[0] = k * a * n1 * sgn(xa) * xan21 * xyp1;
    grad[1] = k * b * n1 * sgn(yb) * ybn21 * xyp1;
    grad[2] = k * c * n1 * sgn(zc) * zcn11;
    if (hess) {
      hess[0] = (k * a) * (n1 * (n2 - 1.0) * xan22 * xyp1 + (n1 - n2) * n1 * xan21 * xan21 * xyp2) * a;
      hess[4] = (k * b) * (n1 * (n2 - 1.0) * ybn22 * xyp1 + (n1 - n2) * n1 * ybn21 * ybn21 * xyp2) * b;
      hess[1] = k * (n1 - n2) * n1 * sgn(xa) * sgn(yb) * (xan21 * a) * (ybn21 * b) * xyp2;
      hess[3] = hess[1];
      hess[8] = (k * c) * (n1 * (n1 - 1.0) * zcn12) * c;
      hess[2] = hess[6] = hess[5] = hess[7] = 0.0;
    }
  }
}

static void matvec3(const double *m, const double *v, double *r) {
  r[0] = m[0] * v[0] + m[1] * v[1] + m[2] * v[2];
  r[1] = m[3] * v[0] + m[4] * v[1] + m[5] * v[2];
  r[2] = m[6] * v[0] + m[7] * v[1] + m[8] * v[2];
}

static void tmatvec3(const double *m, const double *v, double *r) {
  r[0] = m[0] * v[0] + m[3] * v[1] + m[6] * v[2];
  r[1] = m[1] * v[0] + m[4] * v[1] + m[7] * v[2];
  r[2] = m[2] * v[0] + m[5] * v[1] + m[8] * v[2];
}

/* R * M * R^T with the products of times3 / times3_transpose
 * (math_extra_liggghts_nonspherical.h:472-509) */
static void rotate_tensor(const double *R, const double *M, double *out) {
  double t[9];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      t[i * 3 + j] = R[i * 3 + 0] * M[0 * 3 + j] + R[i * 3 + 1] * M[1 * 3 + j] + R[i * 3 + 2] * M[2 * 3 + j];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      out[i * 3 + j] = t[i * 3 + 0] * R[j * 3 + 0] + t[i * 3 + 1] * R[j * 3 + 1] + t[i * 3 + 2] * R[j * 3 + 2];
}

/* shape_function_props_global (superquadric.cpp:313-326) */
static void props_global(const sq_t *s, const double *x, double *f, double *grad, double *hess) {
  double d[3] = {x[0] - s->center[0], x[1] - s->center[1], x[2] - s->center[2]};
  double xl[3], gl[3];
  tmatvec3(s->rot, d, xl);
  if (hess) {
    double hl[9];
    props_local(s, xl, f, gl, hl);
    rotate_tensor(s->rot, hl, hess);
  } else {
    props_local(s, xl, f, gl, NULL);
  }
  matvec3(s->rot, gl, grad);
}

static double dot3(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

static void cross3(const double *a, const double *b, double *r) {
  r[0] = a[1] * b[2] - a[2] * b[1];
  r[1] = a[2] * b[0] - a[0] * b[2];
  r[2] = a[0] * b[1] - a[1] * b[0];
}

static double fmax2(double a, double b) { return a > b ? a : b; }

static double merit_of(const double *gA, const double *gB, double f1, double f2) {
  double v[3];
  cross3(gA, gB, v);
  const double sine = fabs(dot3(v, v) / (dot3(gA, gA) * dot3(gB, gB)));
  const double aaa = fabs(f1 - f2) / (fmax2(fabs(f1), fabs(f2)) + 1e-16);
  return fmax2(sine, aaa * aaa);
}

/* calc_F with a given mu (math_extra_liggghts_superquadric.cpp:619-637) */
static double calc_F_mu(const sq_t *A, const sq_t *B, double *f1, double *f2, double *gA, double *gB,
                        const double *p, double mu, double *F, double *merit) {
  props_global(A, p, f1, gA, NULL);
  props_global(B, p, f2, gB, NULL);
  const double mu_sq = mu * mu;
  for (int k = 0; k < 3; k++) F[k] = gA[k] + mu_sq * gB[k];
  F[3] = *f1 - *f2;
  *merit = merit_of(gA, gB, *f1, *f2);
  return F[0] * F[0] + F[1] * F[1] + F[2] * F[2] + F[3] * F[3];
}

/* calc_F with mu from the gradients (math_extra_liggghts_superquadric.cpp:639-658) */
static double calc_F_free(const sq_t *A, const sq_t *B, double *f1, double *f2, double *gA, double *gB,
                          const double *p, double *mu, double *F, double *merit) {
  props_global(A, p, f1, gA, NULL);
  props_global(B, p, f2, gB, NULL);
  const double mu_sq = -dot3(gA, gB) / dot3(gB, gB);
  *mu = sqrt(fabs(mu_sq));
  for (int k = 0; k < 3; k++) F[k] = gA[k] + mu_sq * gB[k];
  F[3] = *f1 - *f2;
  *merit = merit_of(gA, gB, *f1, *f2);
  return F[0] * F[0] + F[1] * F[1] + F[2] * F[2] + F[3] * F[3];
}

/* determinant_4x4 (math_extra_liggghts_superquadric.cpp:758-775) */
static double det4(const double *m) {
  const double m00 = m[0], m01 = m[1], m02 = m[2], m03 = m[3];
  const double m10 = m[4], m11 = m[5], m12 = m[6], m13 = m[7];
  const double m20 = m[8], m21 = m[9], m22 = m[10], m23 = m[11];
  const double m30 = m[12], m31 = m[13], m32 = m[14], m33 = m[15];
  return m03 * m12 * m21 * m30 - m02 * m13 * m21 * m30 - m03 * m11 * m22 * m30 + m01 * m13 * m22 * m30 +
         m02 * m11 * m23 * m30 - m01 * m12 * m23 * m30 - m03 * m12 * m20 * m31 + m02 * m13 * m20 * m31 +
         m03 * m10 * m22 * m31 - m00 * m13 * m22 * m31 - m02 * m10 * m23 * m31 + m00 * m12 * m23 * m31 +
         m03 * m11 * m20 * m32 - m01 * m13 * m20 * m32 - m03 * m10 * m21 * m32 + m00 * m13 * m21 * m32 +
         m01 * m10 * m23 * m32 - m00 * m11 * m23 * m32 - m02 * m11 * m20 * m33 + m01 * m12 * m20 * m33 +
         m02 * m10 * m21 * m33 - m00 * m12 * m21 * m33 - m01 * m10 * m22 * m33 + m00 * m11 * m22 * m33;
}

/* cofactor of the 4x4 inverse (invf, math_extra_liggghts_superquadric.cpp:715-737) */
static double cofactor4(int i, int j, const double *m) {
  const int o = 2 + (j - i);
  i += 4 + o;
  j += 4 - o;
#define E(a, b) m[((j + (b)) % 4) * 4 + ((i + (a)) % 4)]
  const double inv = +E(+1, -1) * E(+0, +0) * E(-1, +1) + E(+1, +1) * E(+0, -1) * E(-1, +0) +
                     E(-1, -1) * E(+1, +0) * E(+0, +1) - E(-1, -1) * E(+0, +0) * E(+1, +1) -
                     E(-1, +1) * E(+0, -1) * E(+1, +0) - E(+1, -1) * E(-1, +0) * E(+0, +1);
#undef E
  return (o % 2) ? inv : -inv;
}

/* inverseMatrix4x4 (math_extra_liggghts_superquadric.cpp:739-756) */
static double inv4(const double *m, double *out) {
  double inv[16];
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) inv[j * 4 + i] = cofactor4(i, j, m);
  double D = 0;
  for (int k = 0; k < 4; k++) D += m[k] * inv[k * 4];
  if (D == 0) return D;
  const double Di = 1.0 / D;
  for (int i = 0; i < 16; i++) out[i] = inv[i] * Di;
  return D;
}

/* matvecN: zero entries are skipped (math_extra_liggghts_nonspherical.h:384-394) */
static void matvec4(const double *A, const double *b, double *c) {
  for (int i = 0; i < 4; i++) {
    c[i] = 0.0;
    for (int j = 0; j < 4; j++)
      if (A[i * 4 + j] != 0.0 && b[j] != 0.0) c[i] += A[i * 4 + j] * b[j];
  }
}

/* GMRES<4,4> (math_extra_liggghts_nonspherical.h:225-314) */
static void gmres44(const double *A, const double *b, const double *x0, double *x) {
  enum { N = 4, MS = 4 };
  int m = MS;
  double res[N], Ax0[N];
  matvec4(A, x0, Ax0);
  for (int i = 0; i < N; i++) res[i] = b[i] - Ax0[i];
  for (int i = 0; i < N; i++) x[i] = x0[i];
  double beta_sq = 0.0;
  for (int i = 0; i < N; i++) beta_sq += res[i] * res[i];
  if (beta_sq < 1e-20) return;
  const double beta = sqrt(beta_sq);
  double y[MS] = {0}, v[(MS + 1) * N] = {0}, h[(MS + 1) * MS] = {0}, w[MS * N] = {0};
  for (int i = 0; i < N; i++) v[i] = res[i] * (1.0 / beta);
  for (int j = 0; j < m; j++) {
    matvec4(A, v + j * N, w + j * N);
    for (int i = 0; i < j + 1; i++) {
      for (int ix = 0; ix < N; ix++) h[i * m + j] += w[j * N + ix] * v[i * N + ix];
      for (int ix = 0; ix < N; ix++) w[j * N + ix] -= h[i * m + j] * v[i * N + ix];
    }
    double nw = 0.0;
    for (int ix = 0; ix < N; ix++) nw += w[j * N + ix] * w[j * N + ix];
    h[(j + 1) * m + j] = sqrt(nw);
    if (h[(j + 1) * m + j] < 1e-12) {
      m = j + 1;
      break;
    }
    for (int ix = 0; ix < N; ix++) v[(j + 1) * N + ix] = w[j * N + ix] * (1.0 / h[(j + 1) * m + j]);
  }
  double hm[(MS + 1) * MS], g[MS + 1] = {0};
  g[0] = beta;
  for (int i = 0; i < m + 1; i++)
    for (int j = 0; j < m; j++) hm[i * MS + j] = h[i * MS + j];
  double row1[MS], row2[MS];
  for (int j = 0; j < m; j++) {
    const double root_inv = 1.0 / sqrt(hm[j * MS + j] * hm[j * MS + j] + hm[(j + 1) * MS + j] * hm[(j + 1) * MS + j]);
    const double s = hm[(j + 1) * MS + j] * root_inv;
    const double c = hm[j * MS + j] * root_inv;
    for (int i = 0; i < m; i++) {
      row1[i] = c * hm[j * MS + i] + s * hm[(j + 1) * MS + i];
      row2[i] = -s * hm[j * MS + i] + c * hm[(j + 1) * MS + i];
    }
    for (int i = 0; i < m; i++) {
      hm[j * MS + i] = row1[i];
      hm[(j + 1) * MS + i] = row2[i];
    }
    const double g1 = c * g[j] + s * g[j + 1];
    const double g2 = -s * g[j] + c * g[j + 1];
    g[j] = g1;
    g[j + 1] = g2;
  }
  for (int j = 0; j < m; j++) {
    const int j1 = m - 1 - j;
    double summ = 0.0;
    for (int i = 0; i < j; i++) {
      const int i1 = m - 1 - i;
      summ += hm[j1 * MS + i1] * y[i1];
    }
    y[j1] = (g[j1] - summ) / hm[j1 * MS + j1];
    for (int i = 0; i < N; i++) x[i] += y[j1] * v[j1 * N + i];
  }
}

static double min3(double a, double b, double c) {
  double m = a < b ? a : b;
  return m < c ? m : c;
}

/* calc_contact_point (math_extra_liggghts_superquadric.cpp:778-1022); returns
 * the number of Newton iterations, sets *fail */
static int contact_point(sq_t *A, sq_t *B, double ratio, const double *init1, double *result, double *fi,
                         double *fj, int *fail) {
  const double tol1 = 1e-10, tol2 = 1e-12;
  *fail = 0;
  double mu, mu_sq, F[4], F1[4], F2[4], point[3];
  double mu1, mu2, gA1[3], gB1[3], gA2[3], gB2[3];
  double fi1, fj1, fi2, fj2, merit01, merit02, merit0;
  const double res01 = calc_F_free(A, B, &fi1, &fj1, gA1, gB1, init1, &mu1, F1, &merit01);
  if (merit01 < tol1) {
    memcpy(result, init1, 3 * sizeof(double));
    memcpy(A->gradient, gA1, 3 * sizeof(double));
    memcpy(B->gradient, gB1, 3 * sizeof(double));
    *fi = fi1;
    *fj = fj1;
    return 0;
  }
  double init2[3];
  for (int k = 0; k < 3; k++) init2[k] = ratio * B->center[k] + (1.0 - ratio) * A->center[k];
  const double res02 = calc_F_free(A, B, &fi2, &fj2, gA2, gB2, init2, &mu2, F2, &merit02);
  double res0;
  if (res01 < res02) {
    res0 = res01;
    memcpy(A->gradient, gA1, 3 * sizeof(double));
    memcpy(B->gradient, gB1, 3 * sizeof(double));
    memcpy(point, init1, 3 * sizeof(double));
    memcpy(F, F1, 4 * sizeof(double));
    *fi = fi1;
    *fj = fj1;
    mu = mu1;
    merit0 = merit01;
  } else {
    res0 = res02;
    memcpy(A->gradient, gA2, 3 * sizeof(double));
    memcpy(B->gradient, gB2, 3 * sizeof(double));
    memcpy(point, init2, 3 * sizeof(double));
    memcpy(F, F2, 4 * sizeof(double));
    *fi = fi2;
    *fj = fj2;
    mu = mu2;
    merit0 = merit02;
  }
  if (merit0 < tol1) {
    memcpy(result, point, 3 * sizeof(double));
    return 0;
  }
  double merit1 = merit0, merit2 = merit0, res1 = res0, res2 = res1;
  const double size_i = min3(A->shape[0], A->shape[1], A->shape[2]);
  const double size_j = min3(B->shape[0], B->shape[1], B->shape[2]);
  const double size = size_i < size_j ? size_i : size_j;
  double J4[16] = {0}, J4inv[16], delta[4] = {0}, delta_0[4] = {0}, pointa[3], pointb[3];
  const int Niter = 100000;
  int iter;
  for (iter = 0; iter < Niter; iter++) {
    merit2 = merit1;
    res2 = res1;
    {
      double fdummy, gdummy[3];
      props_global(A, point, &fdummy, gdummy, A->hessian);
      props_global(B, point, &fdummy, gdummy, B->hessian);
    }
    mu_sq = mu * mu;
    for (int i = 0; i < 3; i++) {
      for (int j = 0; j < 3; j++) J4[4 * i + j] = A->hessian[i * 3 + j] + mu_sq * B->hessian[i * 3 + j];
      J4[3 + 4 * i] = 2.0 * mu * B->gradient[i];
      J4[i + 4 * 3] = A->gradient[i] - B->gradient[i];
    }
    J4[15] = 0.0;
    const double det = det4(J4);
    if (fabs(det) > 1.0) {
      inv4(J4, J4inv);
      matvec4(J4inv, F, delta);
    } else {
      memcpy(delta_0, delta, sizeof delta);
      gmres44(J4, F, delta_0, delta);
    }
    int nan_found = 0;
    for (int i = 0; i < 4; i++)
      if (isnan(delta[i])) {
        nan_found = 1;
        *fail = 1;
        break;
      }
    int converged = 0;
    const double deltax = fmax2(fmax2(fabs(delta[0]), fabs(delta[1])), fabs(delta[2]));
    if (deltax == 0.0) {
      memcpy(result, point, 3 * sizeof(double));
      converged = 1;
    } else {
      if (nan_found) break;
      double point_[3], mu_, fi_, fj_, merit2_;
      for (int k = 0; k < 3; k++) point_[k] = point[k] - delta[k];
      mu_ = mu - delta[3];
      const double res2_ = calc_F_mu(A, B, &fi_, &fj_, A->gradient, B->gradient, point_, mu_, F, &merit2_);
      if (res2_ < res1 || merit2_ < tol1 || deltax < tol2 * size) {
        merit2 = merit2_;
        res2 = res2_;
        mu = mu_;
        *fi = fi_;
        *fj = fj_;
        memcpy(point, point_, 3 * sizeof(double));
        if (merit2 < tol1 || deltax < tol2 * size) converged = 1;
      } else { /* golden-section search along the Newton step */
        double a = 0.0, b = 1.0, eps = fabs(b - a);
        double res2a = 0, res2b = 0, merit2a = 0, merit2b = 0;
        while (eps > 1e-8) {
          const double alpha1 = b - (b - a) * PHI_INV;
          const double alpha2 = a + (b - a) * PHI_INV;
          for (int k = 0; k < 3; k++) pointa[k] = point[k] + (-alpha1) * delta[k];
          mu1 = mu - alpha1 * delta[3];
          res2a = calc_F_mu(A, B, &fi1, &fj1, gA1, gB1, pointa, mu1, F1, &merit2a);
          for (int k = 0; k < 3; k++) pointb[k] = point[k] + (-alpha2) * delta[k];
          mu2 = mu - alpha2 * delta[3];
          res2b = calc_F_mu(A, B, &fi2, &fj2, gA2, gB2, pointb, mu2, F2, &merit2b);
          if ((res2a < res2b ? res2a : res2b) < res1) {
            if (res2a < res2b) {
              res2 = res2a;
              memcpy(F, F1, 4 * sizeof(double));
              mu = mu1;
              memcpy(A->gradient, gA1, 3 * sizeof(double));
              memcpy(B->gradient, gB1, 3 * sizeof(double));
              memcpy(point, pointa, 3 * sizeof(double));
              *fi = fi1;
              *fj = fj1;
              merit2 = merit2a;
            } else {
              res2 = res2b;
              memcpy(F, F2, 4 * sizeof(double));
              mu = mu2;
              memcpy(A->gradient, gA2, 3 * sizeof(double));
              memcpy(B->gradient, gB2, 3 * sizeof(double));
              memcpy(point, pointb, 3 * sizeof(double));
              *fi = fi2;
              *fj = fj2;
              merit2 = merit2b;
            }
            if (merit2 < tol1) converged = 1;
            eps = 0.0;
          } else {
            if (res2a > res2b)
              a = alpha1;
            else
              b = alpha2;
            eps = fabs(b - a);
          }
        }
        if (eps != 0.0) {
          memcpy(F, F2, 4 * sizeof(double));
          mu = mu2;
          memcpy(A->gradient, gA2, 3 * sizeof(double));
          memcpy(B->gradient, gB2, 3 * sizeof(double));
          memcpy(point, pointb, 3 * sizeof(double));
          *fi = fi2;
          *fj = fj2;
          res2 = res2b;
          merit2 = merit2b;
          if (merit2 < tol1) converged = 1;
        }
      }
      if (fabs(res1 - res2) / fmax2(res1, res2) < 1e-3) converged = 1;
    }
    merit1 = merit2;
    res1 = res2;
    if (iter >= Niter - 1) *fail = 1;
    if (converged || *fail) {
      memcpy(result, point, 3 * sizeof(double));
      break;
    }
  }
  return iter + 1;
}

/* calc_contact_point_if_no_previous_point_avaialable
 * (math_extra_liggghts_superquadric.cpp:1390-1430): N + 1 solves along the
 * homotopy from volume-equivalent spheres to the actual shapes, each
 * starting from the previous one's point */
static int contact_point_fresh(sq_t *A, sq_t *B, double *cp, double *fi, double *fj, int *fail) {
  const double ai = A->shape[0], bi = A->shape[1], ci = A->shape[2];
  const double aj = B->shape[0], bj = B->shape[1], cj = B->shape[2];
  const double ri = cbrt(ai * bi * ci), rj = cbrt(aj * bj * cj);
  const double ratio = ri / (ri + rj);
  const double ni0 = A->blockiness[0], ni1 = A->blockiness[1];
  const double nj0 = B->blockiness[0], nj1 = B->blockiness[1];
  double nmax = fmax2(fmax2(fmax2(ni0, ni1), nj0), nj1);
  const int N1 = (int)((nmax - 2.0) / 1.0 + 1);
  const int N = N1 > 10 ? N1 : 10;
  const double Ninv = 1.0 / (double)N;
  int iters = 0, f = 0;
  *fail = 0;
  for (int i = 0; i <= N; i++) {
    sq_set_blockiness(A, 2.0 + (ni0 - 2.0) * (double)i * Ninv, 2.0 + (ni1 - 2.0) * (double)i * Ninv);
    sq_set_blockiness(B, 2.0 + (nj0 - 2.0) * (double)i * Ninv, 2.0 + (nj1 - 2.0) * (double)i * Ninv);
    sq_set_shape(A, ri + (ai - ri) * (double)i * Ninv, ri + (bi - ri) * (double)i * Ninv,
                 ri + (ci - ri) * (double)i * Ninv);
    sq_set_shape(B, rj + (aj - rj) * (double)i * Ninv, rj + (bj - rj) * (double)i * Ninv,
                 rj + (cj - rj) * (double)i * Ninv);
    if (i == 0)
      for (int k = 0; k < 3; k++) cp[k] = ratio * B->center[k] + (1.0 - ratio) * A->center[k];
    double pre[3] = {cp[0], cp[1], cp[2]};
    iters += contact_point(A, B, ratio, pre, cp, fi, fj, &f);
    *fail = f; /* the last solve's flag, as the reference keeps it */
    sq_set_shape(A, ai, bi, ci);
    sq_set_shape(B, aj, bj, cj);
    sq_set_blockiness(A, ni0, ni1);
    sq_set_blockiness(B, nj0, nj1);
  }
  return iters;
}

/* ---- contact detection around the contact point ---- */

/* shape_function_local (superquadric.cpp:109-136): its own arithmetic, not
 * props_local's */
static double shape_local(const sq_t *s, const double *x) {
  const double n1 = s->blockiness[0], n2 = s->blockiness[1];
  const double xi = x[0] * s->shape_inv[0], yi = x[1] * s->shape_inv[1], zi = x[2] * s->shape_inv[2];
  double f;
  if (s->int_blockiness) {
    const int n1i = (int)floor(n1 + 1e-2), n2i = (int)floor(n2 + 1e-2);
    if (n1i == n2i)
      f = pow_abs_int(xi, n2i) + pow_abs_int(yi, n2i) + pow_abs_int(zi, n1i) - 1.0;
    else
      f = (pow_abs(pow_abs_int(xi, n2i) + pow_abs_int(yi, n2i), n1 / n2) + pow_abs_int(zi, n1i) - 1.0);
  } else {
    f = (pow_abs(pow_abs(xi, n2) + pow_abs(yi, n2), n1 / n2) + pow_abs(zi, n1) - 1.0);
  }
  return s->koef * f;
}

/* shape_function_gradient_local (superquadric.cpp:139-176) */
static void gradient_local(const sq_t *s, const double *x, double *r) {
  const double n1 = s->blockiness[0], n2 = s->blockiness[1];
  const double a = s->shape_inv[0], b = s->shape_inv[1], c = s->shape_inv[2];
  const double xa = x[0] * a, yb = x[1] * b, zc = x[2] * c;
  double xan21, ybn21, zcn11;
  if (s->int_blockiness) {
    const int n1i = (int)floor(n1 + 1e-2), n2i = (int)floor(n2 + 1e-2);
    xan21 = pow_abs_int(xa, n2i - 1);
    ybn21 = pow_abs_int(yb, n2i - 1);
    zcn11 = pow_abs_int(zc, n1i - 1);
  } else {
    xan21 = pow_abs(xa, n2 - 1.0);
    ybn21 = pow_abs(yb, n2 - 1.0);
    zcn11 = pow_abs(zc, n1 - 1.0);
  }
  const double k = s->koef;
  if (comp_double(n1, n2, 1e-2)) {
    r[0] = n1 * (k * a) * sgn(xa) * xan21;
    r[1] = n1 * (k * b) * sgn(yb) * ybn21;
    r[2] = n1 * (k * c) * sgn(zc) * zcn11;
  } else {
    const double xy = xan21 * fabs(xa) + ybn21 * fabs(yb);
    const double xyp1 = pow_abs(xy, n1 / n2 - 1.0);
    r[0] = n1 * (k * a) * sgn(xa) * xan21 * xyp1;
    r[1] = n1 * (k * b) * sgn(yb) * ybn21 * xyp1;
    r[2] = n1 * (k * c) * sgn(zc) * zcn11;
  }
}

static double clamp(double v, double lo, double hi) { return v < lo ? lo : (v > hi ? hi : v); }

/* Superquadric::surface_line_intersection (superquadric.cpp:712-816) with
 * ellipsoid_line_intersection_local (:1011-1035): the point where the line
 * start + alpha * dir crosses the surface, alpha returned */
static double surface_line_intersection(const sq_t *s, int max_iters, int use_alpha_ellipsoid,
                                        const double *start, const double *dir, double alpha_prev,
                                        double *result) {
  double sl[3], dl[3], point[3], grad[3];
  const double d0[3] = {start[0] - s->center[0], start[1] - s->center[1], start[2] - s->center[2]};
  tmatvec3(s->rot, d0, sl);
  tmatvec3(s->rot, dir, dl);
  double alpha = alpha_prev;
  const double eps0 = 1e-14;
  double alpha_ell = 0.0;
  const int ellipsoid = comp_double(s->blockiness[0], 2.0, 1e-2) && comp_double(s->blockiness[1], 2.0, 1e-2);
  if (ellipsoid || use_alpha_ellipsoid) {
    const double x0 = sl[0] * s->shape_inv[0], nx = dl[0] * s->shape_inv[0];
    const double y0 = sl[1] * s->shape_inv[1], ny = dl[1] * s->shape_inv[1];
    const double z0 = sl[2] * s->shape_inv[2], nz = dl[2] * s->shape_inv[2];
    const double qa = nx * nx + ny * ny + nz * nz;
    const double qb = 2.0 * (x0 * nx + y0 * ny + z0 * nz);
    const double qc = x0 * x0 + y0 * y0 + z0 * z0 - 1.0;
    const double D = qb * qb - 4.0 * qa * qc;
    const double two_a_inv = 0.5 / qa;
    double a1, a2;
    if (D < 0) {
      a1 = -qb * two_a_inv;
      a2 = a1;
    } else {
      const double sq = sqrt(D);
      a1 = (-qb + sq) * two_a_inv;
      a2 = (-qb - sq) * two_a_inv;
    }
    if (ellipsoid) {
      if (a1 > -eps0 && a2 < -eps0)
        alpha = a1;
      else if (a1 < -eps0 && a2 > -eps0)
        alpha = a2;
      else if (a1 < -eps0 && a2 < -eps0)
        alpha = fmax2(a1, a2);
      else
        alpha = a1 < a2 ? a1 : a2;
    } else {
      const double f = shape_local(s, sl);
      if (f > eps0)
        alpha_ell = a1 < a2 ? a1 : a2;
      else if (f < -eps0)
        alpha_ell = fmax2(a1, a2);
      else
        alpha_ell = alpha = alpha_prev = 0.0;
    }
  }
  if (!ellipsoid) {
    const double a = 0.5 * min3(s->shape[0], s->shape[1], s->shape[2]) /
                     sqrt(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]);
    const double eps = 1e-16;
    alpha = clamp(alpha, -a, a);
    for (int k = 0; k < 3; k++) point[k] = sl[k] + alpha * dl[k];
    double f0 = shape_local(s, point);
    if (use_alpha_ellipsoid) {
      double pe[3];
      for (int k = 0; k < 3; k++) pe[k] = sl[k] + alpha_ell * dl[k];
      const double f1 = shape_local(s, pe);
      if (fabs(f1) < fabs(f0)) {
        alpha = alpha_ell;
        f0 = f1;
        memcpy(point, pe, sizeof pe);
      }
    }
    if (fabs(f0) < eps) {
      double g[3];
      matvec3(s->rot, point, g);
      for (int k = 0; k < 3; k++) result[k] = g[k] + s->center[k];
      return alpha;
    }
    for (int i = 0; i < max_iters; i++) {
      if (fabs(f0) < eps * s->koef) break;
      gradient_local(s, point, grad);
      const double f_der = dot3(grad, dl);
      double delta;
      if (fabs(f_der) < 1e-10) {
        delta = 1e-10;
        double p_[3], g_[3];
        for (int k = 0; k < 3; k++) p_[k] = point[k] + delta * dl[k];
        gradient_local(s, p_, g_);
        if (dot3(g_, dl) < 0.0) delta = -delta;
      } else {
        delta = clamp(-f0 / f_der, -a, a);
      }
      if (fabs(delta) <= eps) {
        alpha += delta;
        for (int k = 0; k < 3; k++) point[k] = point[k] + delta * dl[k];
        break;
      }
      while (fabs(delta) > eps) {
        double p_[3];
        for (int k = 0; k < 3; k++) p_[k] = point[k] + delta * dl[k];
        const double f_ = shape_local(s, p_);
        if (fabs(f_) < fabs(f0) || fabs(f_) < eps * s->koef) {
          f0 = f_;
          alpha += delta;
          memcpy(point, p_, sizeof p_);
          break;
        }
        delta *= 0.5;
      }
    }
  }
  for (int k = 0; k < 3; k++) result[k] = start[k] + alpha * dir[k];
  return alpha;
}

/* check_inequalities (math_extra_liggghts_superquadric.cpp:72-150): axis i
 * separates the two boxes */
static int obb_separated(int i, const double *a, const double *b, const double *Ad, const double *Bd,
                         const double *C) {
  double R = 0.0, R0 = 0.0, R1 = 0.0;
  switch (i) {
    case 0: R0 = a[0]; R1 = b[0] * fabs(C[0]) + b[1] * fabs(C[1]) + b[2] * fabs(C[2]); R = fabs(Ad[0]); break;
    case 1: R0 = a[1]; R1 = b[0] * fabs(C[3]) + b[1] * fabs(C[4]) + b[2] * fabs(C[5]); R = fabs(Ad[1]); break;
    case 2: R0 = a[2]; R1 = b[0] * fabs(C[6]) + b[1] * fabs(C[7]) + b[2] * fabs(C[8]); R = fabs(Ad[2]); break;
    case 3: R0 = a[0] * fabs(C[0]) + a[1] * fabs(C[3]) + a[2] * fabs(C[6]); R1 = b[0]; R = fabs(Bd[0]); break;
    case 4: R0 = a[0] * fabs(C[1]) + a[1] * fabs(C[4]) + a[2] * fabs(C[7]); R1 = b[1]; R = fabs(Bd[1]); break;
    case 5: R0 = a[0] * fabs(C[2]) + a[1] * fabs(C[5]) + a[2] * fabs(C[8]); R1 = b[2]; R = fabs(Bd[2]); break;
    case 6: R0 = a[1] * fabs(C[6]) + a[2] * fabs(C[3]); R1 = b[1] * fabs(C[2]) + b[2] * fabs(C[1]); R = fabs(C[3] * Ad[2] - C[6] * Ad[1]); break;
    case 7: R0 = a[1] * fabs(C[7]) + a[2] * fabs(C[4]); R1 = b[0] * fabs(C[2]) + b[2] * fabs(C[0]); R = fabs(C[4] * Ad[2] - C[7] * Ad[1]); break;
    case 8: R0 = a[1] * fabs(C[8]) + a[2] * fabs(C[5]); R1 = b[0] * fabs(C[1]) + b[1] * fabs(C[0]); R = fabs(C[5] * Ad[2] - C[8] * Ad[1]); break;
    case 9: R0 = a[0] * fabs(C[6]) + a[2] * fabs(C[0]); R1 = b[1] * fabs(C[5]) + b[2] * fabs(C[4]); R = fabs(C[6] * Ad[0] - C[0] * Ad[2]); break;
    case 10: R0 = a[0] * fabs(C[7]) + a[2] * fabs(C[1]); R1 = b[0] * fabs(C[5]) + b[2] * fabs(C[3]); R = fabs(C[7] * Ad[0] - C[1] * Ad[2]); break;
    case 11: R0 = a[0] * fabs(C[8]) + a[2] * fabs(C[2]); R1 = b[0] * fabs(C[4]) + b[1] * fabs(C[3]); R = fabs(C[8] * Ad[0] - C[2] * Ad[2]); break;
    case 12: R0 = a[0] * fabs(C[3]) + a[1] * fabs(C[0]); R1 = b[1] * fabs(C[8]) + b[2] * fabs(C[7]); R = fabs(C[0] * Ad[1] - C[3] * Ad[0]); break;
    case 13: R0 = a[0] * fabs(C[4]) + a[1] * fabs(C[1]); R1 = b[0] * fabs(C[8]) + b[2] * fabs(C[6]); R = fabs(C[1] * Ad[1] - C[4] * Ad[0]); break;
    case 14: R0 = a[0] * fabs(C[5]) + a[1] * fabs(C[2]); R1 = b[0] * fabs(C[7]) + b[1] * fabs(C[6]); R = fabs(C[2] * Ad[1] - C[5] * Ad[0]); break;
  }
  if (fabs(R0 + R1) < 1e-14 && fabs(R) < 1e-14) return 0;
  return R > R0 + R1;
}

/* obb_intersect (math_extra_liggghts_superquadric.cpp:166-203): the boxes of
 * the half-axes grown by scale_koef = 1.0001 (math_extra_liggghts_nonspherical.h:60);
 * the last separating axis found is tried first next time */
static int obb_intersect(const sq_t *A, const sq_t *B, int *start) {
  const double scale_koef = 1.0001;
  const double delta[3] = {A->center[0] - B->center[0], A->center[1] - B->center[1], A->center[2] - B->center[2]};
  double a[3], b[3];
  for (int k = 0; k < 3; k++) {
    a[k] = A->shape[k] * scale_koef;
    b[k] = B->shape[k] * scale_koef;
  }
  const double *m = A->rot, *m2 = B->rot;
  double C[9];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) C[i * 3 + j] = m[0 * 3 + i] * m2[0 * 3 + j] + m[1 * 3 + i] * m2[1 * 3 + j] + m[2 * 3 + i] * m2[2 * 3 + j];
  double Ad[3], Bd[3];
  tmatvec3(A->rot, delta, Ad);
  tmatvec3(B->rot, delta, Bd);
  for (int i = 0; i < 15; i++) {
    const int which = i == 0 ? *start : (i == *start ? 0 : i);
    if (obb_separated(which, a, b, Ad, Bd, C)) {
      *start = which;
      return 0;
    }
  }
  return 1;
}

/* SurfaceModel<SURFACE_SUPERQUADRIC>::checkSurfaceIntersect
 * (surface_model_superquadric.h:108-205) for one pair, with the contact
 * history (state, first OBB inequality, contact point, alpha_i, alpha_j) in
 * hist[7]; calc_contact_point_using_prev_step (math_extra_liggghts_superquadric.cpp:
 * 1434-1466) takes the pair's velocities (vel: v_i, v_j, omega_i, omega_j),
 * masses and dt. Returns 1 when the surfaces intersect. */
int orc_sq_check(const orc_sq_pose *pa, const orc_sq_pose *pb, double *hist, const double *vel,
                 const double *mass, double dt, double *en, double *deltan, double *cpa, double *cpb,
                 double *delta, int *iters) {
  enum { FAR = 0, CLOSE = 1, INTERSECT = 2 };
  sq_t A, B;
  sq_init(&A, pa);
  sq_init(&B, pb);
  int start = (int)(hist[1] + 0.5);
  int obb;
  const int state = (int)(hist[0] + 0.5);
  if (state == INTERSECT)
    obb = 1;
  else
    obb = obb_intersect(&A, &B, &start);
  int in_contact = 0;
  *iters = 0;
  if (obb) {
    double fi = 0, fj = 0, cp[3];
    int fail = 0;
    const double ri = cbrt(A.shape[0] * A.shape[1] * A.shape[2]);
    const double rj = cbrt(B.shape[0] * B.shape[1] * B.shape[2]);
    const double ratio = ri / (ri + rj);
    if (state == FAR) {
      *iters = contact_point_fresh(&A, &B, cp, &fi, &fj, &fail);
    } else {
      const double *prev = hist + 2;
      double est[3];
      if (vel && mass) {
        const double *vi = vel, *vj = vel + 3, *wi = vel + 6, *wj = vel + 9;
        double xci[3], xcj[3], ri_[3], rj_[3];
        for (int k = 0; k < 3; k++) {
          xci[k] = prev[k] - A.center[k];
          xcj[k] = prev[k] - B.center[k];
        }
        cross3(wi, xci, ri_);
        cross3(wj, xcj, rj_);
        for (int k = 0; k < 3; k++) {
          const double vik = vi[k] + ri_[k], vjk = vj[k] + rj_[k];
          const double veff = (vik * mass[0] + vjk * mass[1]) / (mass[0] + mass[1]);
          est[k] = prev[k] + veff * dt;
        }
      } else {
        memcpy(est, prev, sizeof est);
      }
      *iters = contact_point(&A, &B, ratio, est, cp, &fi, &fj, &fail);
    }
    memcpy(hist + 2, cp, sizeof cp);
    in_contact = (fi > fj ? fi : fj) < 0.0;
    if (in_contact) {
      for (int k = 0; k < 3; k++) en[k] = B.gradient[k] - A.gradient[k];
      const double nrm = sqrt(en[0] * en[0] + en[1] * en[1] + en[2] * en[2]);
      const double inv = nrm == 0. ? 0. : 1. / nrm;
      for (int k = 0; k < 3; k++) en[k] *= inv;
      /* extended_overlap_algorithm (math_extra_liggghts_superquadric.cpp:1489-1504) */
      const double men[3] = {-en[0], -en[1], -en[2]};
      hist[5] = surface_line_intersection(&A, 20, 0, cp, men, hist[5], cpa);
      hist[6] = surface_line_intersection(&B, 20, 0, cp, en, hist[6], cpb);
      for (int k = 0; k < 3; k++) delta[k] = cpb[k] - cpa[k];
      *deltan = fabs(hist[5]) + fabs(hist[6]);
      hist[0] = INTERSECT;
    } else {
      hist[0] = CLOSE;
    }
  } else {
    hist[0] = FAR;
  }
  hist[1] = (double)start;
  return in_contact;
}

void orc_sq_check_batch(int64_t n, const orc_sq_pose *a, const orc_sq_pose *b, double *hist, const double *vel,
                        const double *mass, double dt, int *contact, double *en, double *deltan, double *cpa,
                        double *cpb, int *iters) {
  for (int64_t k = 0; k < n; k++) {
    double d[3];
    contact[k] = orc_sq_check(&a[k], &b[k], hist + 7 * k, vel ? vel + 12 * k : NULL, mass ? mass + 2 * k : NULL,
                              dt, en + 3 * k, deltan + k, cpa + 3 * k, cpb + 3 * k, d, iters + k);
  }
}

int orc_sq_obb(const orc_sq_pose *pa, const orc_sq_pose *pb, int *start) {
  sq_t A, B;
  sq_init(&A, pa);
  sq_init(&B, pb);
  return obb_intersect(&A, &B, start);
}

double orc_sq_line(const orc_sq_pose *p, const double *start, const double *dir, double alpha_prev,
                   double *result) {
  sq_t s;
  sq_init(&s, p);
  return surface_line_intersection(&s, 20, 0, start, dir, alpha_prev, result);
}

/* Superquadric::plane_intersection (superquadric.cpp:440-507): normal and
 * x0 as passed by the wall (un-normalised normal; the 1e-10 tests are on its
 * raw components). Returns 1 when particle and plane intersect. */
int orc_sq_plane(const orc_sq_pose *p, const double *normal_global, const double *x0_global, double *pmax,
                 double *plow) {
  sq_t s;
  sq_init(&s, p);
  double nl[3], x0l[3], pm[3], pl[3];
  tmatvec3(s.rot, normal_global, nl);
  const double d0[3] = {x0_global[0] - s.center[0], x0_global[1] - s.center[1], x0_global[2] - s.center[2]};
  tmatvec3(s.rot, d0, x0l);
  if (dot3(nl, x0l) < 0.0) {
    nl[0] = -nl[0];
    nl[1] = -nl[1];
    nl[2] = -nl[2];
  }
  const double n1 = s.blockiness[0], n2 = s.blockiness[1];
  const double nx = nl[0], ny = nl[1], nz = nl[2];
  const double a = s.shape[0], b = s.shape[1], c = s.shape[2];
  const double d = dot3(nl, x0l);
  if (comp_double(nx, 0.0, 1e-10) && comp_double(ny, 0.0, 1e-10)) {
    pm[0] = pm[1] = 0.0;
    pm[2] = 1.0;
    pl[0] = pl[1] = 0.0;
    pl[2] = d / nz;
  } else {
    if (fabs(nx) > fabs(ny)) {
      const double alpha = pow_abs((ny * b) / (nx * a), 1.0 / (n2 - 1.0));
      const double gamma1 = 1.0 + pow_abs(alpha, n2);
      const double gamma = pow_abs(gamma1, n1 / n2 - 1.0);
      const double beta = pow_abs((nz * c) / (nx * a) * gamma, 1.0 / (n1 - 1.0));
      pm[0] = 1.0 / pow_abs(pow_abs(gamma1, n1 / n2) + pow_abs(beta, n1), 1.0 / n1);
      pm[1] = alpha * pm[0];
      pm[2] = beta * pm[0];
      pl[0] = fabs(d) / (fabs(nx) * a + alpha * fabs(ny) * b + beta * fabs(nz) * c);
      pl[1] = alpha * pl[0];
      pl[2] = beta * pl[0];
    } else {
      const double alpha = pow_abs((nx * a) / (ny * b), 1.0 / (n2 - 1.0));
      const double gamma1 = 1.0 + pow_abs(alpha, n2);
      const double gamma = pow_abs(gamma1, n1 / n2 - 1.0);
      const double beta = pow_abs((nz * c) / (ny * b) * gamma, 1.0 / (n1 - 1.0));
      pm[1] = 1.0 / pow_abs(pow_abs(gamma1, n1 / n2) + pow_abs(beta, n1), 1.0 / n1);
      pm[0] = alpha * pm[1];
      pm[2] = beta * pm[1];
      pl[1] = fabs(d) / (fabs(ny) * b + alpha * fabs(nx) * a + beta * fabs(nz) * c);
      pl[0] = alpha * pl[1];
      pl[2] = beta * pl[1];
    }
    for (int i = 0; i < 3; i++) pl[i] *= s.shape[i] * sgn(nl[i]);
  }
  for (int i = 0; i < 3; i++) pm[i] *= s.shape[i] * sgn(nl[i]);
  const double ff = shape_local(&s, pl);
  double g[3];
  matvec3(s.rot, pm, g);
  for (int k = 0; k < 3; k++) pmax[k] = g[k] + s.center[k];
  matvec3(s.rot, pl, g);
  for (int k = 0; k < 3; k++) plow[k] = g[k] + s.center[k];
  return ff < 0.0;
}

/* bounding_sphere_radius_superquadric (math_extra_liggghts_superquadric.cpp:
 * 206-236), grown by scale_koef = 1.0001 */
double orc_sq_bounding_radius(const double *shape, const double *blk) {
  double r;
  if (comp_double(blk[0], 2.0, 1e-13) && comp_double(blk[1], 2.0, 1e-13)) {
    r = fmax2(fmax2(shape[0], shape[1]), shape[2]);
  } else {
    const double a = fmax2(shape[0], shape[1]);
    const double b = shape[0] < shape[1] ? shape[0] : shape[1];
    const double c = shape[2];
    if (comp_double(blk[1], 2.0, 1e-13)) {
      r = sqrt(a * a + c * c);
    } else {
      const double n1 = fmax2(2.1, blk[0]);
      const double n2 = fmax2(2.1, blk[1]);
      const double alpha = pow_abs(b / a, 2.0 / (n2 - 2.0));
      const double gamma = pow_abs(1.0 + pow_abs(alpha, n2), n1 / n2 - 1.0);
      const double beta = pow_abs(c * c / (a * a) * gamma, 1.0 / (n1 - 2.0));
      const double x0 = 1.0 / pow_abs(pow_abs(1.0 + pow_abs(alpha, n2), n1 / n2) + pow_abs(beta, n1), 1.0 / n1);
      const double xt = x0 * a, yt = alpha * x0 * b, zt = beta * x0 * c;
      r = sqrt(xt * xt + yt * yt + zt * zt);
    }
  }
  return r * 1.0001;
}

void orc_sq_props(const orc_sq_pose *p, const double *x, double *f, double *grad, double *hess) {
  sq_t s;
  sq_init(&s, p);
  props_global(&s, x, f, grad, hess);
}

int orc_sq_contact(const orc_sq_pose *pa, const orc_sq_pose *pb, const double *init, double *point,
                   double *fi, double *fj, int *iters) {
  sq_t A, B;
  sq_init(&A, pa);
  sq_init(&B, pb);
  int fail = 0;
  if (init) {
    const double ri = cbrt(A.shape[0] * A.shape[1] * A.shape[2]);
    const double rj = cbrt(B.shape[0] * B.shape[1] * B.shape[2]);
    *iters = contact_point(&A, &B, ri / (ri + rj), init, point, fi, fj, &fail);
  } else {
    *iters = contact_point_fresh(&A, &B, point, fi, fj, &fail);
  }
  return fail;
}

void orc_sq_contact_batch(int64_t n, const orc_sq_pose *pa, const orc_sq_pose *pb, const double *init,
                          double *point, double *fi, double *fj, int *iters, int *fail) {
  for (int64_t k = 0; k < n; k++)
    fail[k] = orc_sq_contact(&pa[k], &pb[k], init ? init + 3 * k : NULL, point + 3 * k, fi + k, fj + k,
                             iters + k);
}
