// Superquadric contact-point solve on the device, FP64, one thread per pair.
// Same arithmetic and control flow as the reference's
//   Superquadric::shape_function_props_local/_global (superquadric.cpp:237-326),
//   calc_F (math_extra_liggghts_superquadric.cpp:619-658),
//   determinant_4x4 / inverseMatrix4x4 (:715-775),
//   GMRES<4,4> (math_extra_liggghts_nonspherical.h:225-314),
//   calc_contact_point (math_extra_liggghts_superquadric.cpp:778-1022) and the
//   homotopy of calc_contact_point_if_no_previous_point_avaialable (:1390-1430),
// written for registers: fixed-size arrays, no heap, no recursion.
#pragma once

#include <hip/hip_runtime.h>

namespace sqd {

constexpr double kPhiInv = 0.61803398874989479;

struct Sq {
  double center[3], rot[9];
  double shape[3], shape_inv[3];
  double blk[2], koef;
  int int_blk;
  double gradient[3], hessian[9];
};

__device__ __forceinline__ int comp_double(double a, double b, double prec) {
  if (a == b) return 1;
  if (b == 0) return a < prec && a > -prec;
  const double x = a - b;
  return x < prec && x > -prec;
}

__device__ __forceinline__ double pow_abs_int(double a, int b) {
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

// pow_abs_int for the exponents 0..9 without branches: every candidate
// product is formed (a handful of multiplications, in the same order as the
// switch above) and one is selected, so independent calls can overlap
__device__ __forceinline__ double pow_abs_int_sel(double a, int b) {
  a = fabs(a);
  if (b > 9 || b < 0) return pow_abs_int(a, b);
  const double a2 = a * a, a3 = a2 * a, a4 = a2 * a2, a6 = a3 * a3;
  const double a5 = a4 * a, a7 = a4 * a3, a8 = a4 * a4, a9 = a6 * a3;
  double r = 1.0;
  r = b == 1 ? a : r;
  r = b == 2 ? a2 : r;
  r = b == 3 ? a3 : r;
  r = b == 4 ? a4 : r;
  r = b == 5 ? a5 : r;
  r = b == 6 ? a6 : r;
  r = b == 7 ? a7 : r;
  r = b == 8 ? a8 : r;
  r = b == 9 ? a9 : r;
  return r;
}

// the same value as the branching form (MathExtraLiggghtsNonspherical::
// pow_abs): every branch's result is computed and the one the branches would
// take is selected, so the shape function's independent powers (three per
// particle, two particles per Newton step) run as overlapping straight-line
// chains instead of one after another
__device__ __forceinline__ double pow_abs(double base, double exponent) {
  base = fabs(base);
  const double tol = 1e-12;
  const double ae = fabs(exponent);
  const double n = floor(ae);
  const double r0 = pow_abs_int_sel(base, (int)n);
  const double nd = ae - n;
  const double pw = pow(base, nd);
  const double sq = sqrt(base);
  double r = nd > 0.0 ? r0 * (fabs(2.0 * nd - 1.0) < 2.0 * tol ? sq : pw) : r0;
  r = exponent > 0.0 ? r : 1.0 / r;
  r = fabs(base) < tol ? 0.0 : r;
  r = (ae < tol || fabs(base - 1.0) < tol) ? 1.0 : r;
  return r;
}

__device__ __forceinline__ double sgn(double v) { return v > 1e-14 ? 1.0 : (v < -1e-14 ? -1.0 : 0.0); }

__device__ __forceinline__ int is_integer(double x) { return comp_double(x, floor(x + 1e-2), 1e-2); }

__device__ __forceinline__ void set_blockiness(Sq &s, double n1, double n2) {
  s.blk[0] = n1;
  s.blk[1] = n2;
  s.koef = pow_abs(0.5, (n1 > n2 ? n1 : n2) - 2.0);
  const int ellipsoid = comp_double(n1, 2.0, 1e-2) && comp_double(n2, 2.0, 1e-2);
  s.int_blk = ellipsoid || (is_integer(n1) && is_integer(n2));
}

__device__ __forceinline__ void set_shape(Sq &s, double a, double b, double c) {
  s.shape[0] = a;
  s.shape[1] = b;
  s.shape[2] = c;
  s.shape_inv[0] = 1.0 / a;
  s.shape_inv[1] = 1.0 / b;
  s.shape_inv[2] = 1.0 / c;
}

// pose: center[3], quat[4] (w,i,j,k), shape[3], blockiness[2]
__device__ __forceinline__ void init(Sq &s, const double *pose) {
  const double *q = pose + 3;
  const double w2 = q[0] * q[0], i2 = q[1] * q[1], j2 = q[2] * q[2], k2 = q[3] * q[3];
  const double twoij = 2.0 * q[1] * q[2], twoik = 2.0 * q[1] * q[3], twojk = 2.0 * q[2] * q[3];
  const double twoiw = 2.0 * q[1] * q[0], twojw = 2.0 * q[2] * q[0], twokw = 2.0 * q[3] * q[0];
  s.rot[0] = w2 + i2 - j2 - k2;
  s.rot[1] = twoij - twokw;
  s.rot[2] = twojw + twoik;
  s.rot[3] = twoij + twokw;
  s.rot[4] = w2 - i2 + j2 - k2;
  s.rot[5] = twojk - twoiw;
  s.rot[6] = twoik - twojw;
  s.rot[7] = twojk + twoiw;
  s.rot[8] = w2 - i2 - j2 + k2;
  for (int k = 0; k < 3; k++) s.center[k] = pose[k];
  set_shape(s, pose[7], pose[8], pose[9]);
  set_blockiness(s, pose[10], pose[11]);
  for (int k = 0; k < 3; k++) s.gradient[k] = 0.0;
  for (int k = 0; k < 9; k++) s.hessian[k] = 0.0;
}

template <bool HESS>
__device__ __forceinline__ void props_local(const Sq &s, const double *x, double &f, double *grad,
                                            double *hess) {
  const double n1 = s.blk[0], n2 = s.blk[1];
  const double a = s.shape_inv[0], b = s.shape_inv[1], c = s.shape_inv[2];
  const double xa = x[0] * a, yb = x[1] * b, zc = x[2] * c;
  double xan22, ybn22, zcn12;
  if (s.int_blk) {
    const int n1i = (int)floor(n1 + 1e-2), n2i = (int)floor(n2 + 1e-2);
    xan22 = pow_abs_int_sel(xa, n2i - 2);
    ybn22 = pow_abs_int_sel(yb, n2i - 2);
    zcn12 = pow_abs_int_sel(zc, n1i - 2);
  } else {
    xan22 = pow_abs(xa, n2 - 2.0);
    ybn22 = pow_abs(yb, n2 - 2.0);
    zcn12 = pow_abs(zc, n1 - 2.0);
  }
  const double xan21 = xan22 * fabs(xa), ybn21 = ybn22 * fabs(yb), zcn11 = zcn12 * fabs(zc);
  const double xan2 = xan21 * fabs(xa), ybn2 = ybn21 * fabs(yb), zcn1 = zcn11 * fabs(zc);
  const double k = s.koef;
  if (comp_double(n1, n2, 1e-2)) {
    f = k * (xan2 + ybn2 + zcn1 - 1.0);
    grad[0] = k * a * n1 * sgn(xa) * xan21;
    grad[1] = k * b * n1 * sgn(yb) * ybn21;
    grad[2] = k * c * n1 * sgn(zc) * zcn11;
    if (HESS) {
      hess[0] = (k * a) * (n1 * (n2 - 1.0) * xan22) * a;
      hess[4] = (k * b) * (n1 * (n2 - 1.0) * ybn22) * b;
      hess[8] = (k * c) * (n1 * (n1 - 1.0) * zcn12) * c;
      hess[1] = hess[3] = hess[2] = hess[6] = hess[5] = hess[7] = 0.0;
    }
  } else {
    f = k * (pow_abs(xan2 + ybn2, n1 / n2) + zcn1 - 1.0);
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
    if (HESS) {
      hess[0] = (k * a) * (n1 * (n2 - 1.0) * xan22 * xyp1 + (n1 - n2) * n1 * xan21 * xan21 * xyp2) * a;
      hess[4] = (k * b) * (n1 * (n2 - 1.0) * ybn22 * xyp1 + (n1 - n2) * n1 * ybn21 * ybn21 * xyp2) * b;
      hess[1] = k * (n1 - n2) * n1 * sgn(xa) * sgn(yb) * (xan21 * a) * (ybn21 * b) * xyp2;
      hess[3] = hess[1];
      hess[8] = (k * c) * (n1 * (n1 - 1.0) * zcn12) * c;
      hess[2] = hess[6] = hess[5] = hess[7] = 0.0;
    }
  }
}

template <bool HESS>
__device__ __forceinline__ void props_global(const Sq &s, const double *x, double &f, double *grad,
                                             double *hess) {
  const double d[3] = {x[0] - s.center[0], x[1] - s.center[1], x[2] - s.center[2]};
  const double *R = s.rot;
  const double xl[3] = {R[0] * d[0] + R[3] * d[1] + R[6] * d[2], R[1] * d[0] + R[4] * d[1] + R[7] * d[2],
                        R[2] * d[0] + R[5] * d[1] + R[8] * d[2]};
  double gl[3];
  if (HESS) {
    double hl[9], t[9];
    props_local<true>(s, xl, f, gl, hl);
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++)
        t[i * 3 + j] = R[i * 3 + 0] * hl[0 * 3 + j] + R[i * 3 + 1] * hl[1 * 3 + j] + R[i * 3 + 2] * hl[2 * 3 + j];
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++)
        hess[i * 3 + j] = t[i * 3 + 0] * R[j * 3 + 0] + t[i * 3 + 1] * R[j * 3 + 1] + t[i * 3 + 2] * R[j * 3 + 2];
  } else {
    props_local<false>(s, xl, f, gl, nullptr);
  }
  grad[0] = R[0] * gl[0] + R[1] * gl[1] + R[2] * gl[2];
  grad[1] = R[3] * gl[0] + R[4] * gl[1] + R[5] * gl[2];
  grad[2] = R[6] * gl[0] + R[7] * gl[1] + R[8] * gl[2];
}

__device__ __forceinline__ double dot3(const double *a, const double *b) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

__device__ __forceinline__ double merit_of(const double *gA, const double *gB, double f1, double f2) {
  const double v[3] = {gA[1] * gB[2] - gA[2] * gB[1], gA[2] * gB[0] - gA[0] * gB[2],
                       gA[0] * gB[1] - gA[1] * gB[0]};
  const double sine = fabs(dot3(v, v) / (dot3(gA, gA) * dot3(gB, gB)));
  const double aaa = fabs(f1 - f2) / (fmax(fabs(f1), fabs(f2)) + 1e-16);
  return fmax(sine, aaa * aaa);
}

__device__ __forceinline__ double calc_F_mu(const Sq &A, const Sq &B, double &f1, double &f2, double *gA,
                                            double *gB, const double *p, double mu, double *F,
                                            double &merit) {
  props_global<false>(A, p, f1, gA, nullptr);
  props_global<false>(B, p, f2, gB, nullptr);
  const double mu_sq = mu * mu;
  for (int k = 0; k < 3; k++) F[k] = gA[k] + mu_sq * gB[k];
  F[3] = f1 - f2;
  merit = merit_of(gA, gB, f1, f2);
  return F[0] * F[0] + F[1] * F[1] + F[2] * F[2] + F[3] * F[3];
}

__device__ __forceinline__ double calc_F_free(const Sq &A, const Sq &B, double &f1, double &f2, double *gA,
                                              double *gB, const double *p, double &mu, double *F,
                                              double &merit) {
  props_global<false>(A, p, f1, gA, nullptr);
  props_global<false>(B, p, f2, gB, nullptr);
  const double mu_sq = -dot3(gA, gB) / dot3(gB, gB);
  mu = sqrt(fabs(mu_sq));
  for (int k = 0; k < 3; k++) F[k] = gA[k] + mu_sq * gB[k];
  F[3] = f1 - f2;
  merit = merit_of(gA, gB, f1, f2);
  return F[0] * F[0] + F[1] * F[1] + F[2] * F[2] + F[3] * F[3];
}

__device__ __forceinline__ double det4(const double *m) {
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

__device__ __forceinline__ double cofactor4(int i, int j, const double *m) {
  const int o = 2 + (j - i);
  i += 4 + o;
  j += 4 - o;
#define SQD_E(a, b) m[((j + (b)) % 4) * 4 + ((i + (a)) % 4)]
  const double inv = +SQD_E(+1, -1) * SQD_E(+0, +0) * SQD_E(-1, +1) + SQD_E(+1, +1) * SQD_E(+0, -1) * SQD_E(-1, +0) +
                     SQD_E(-1, -1) * SQD_E(+1, +0) * SQD_E(+0, +1) - SQD_E(-1, -1) * SQD_E(+0, +0) * SQD_E(+1, +1) -
                     SQD_E(-1, +1) * SQD_E(+0, -1) * SQD_E(+1, +0) - SQD_E(+1, -1) * SQD_E(-1, +0) * SQD_E(+0, +1);
#undef SQD_E
  return (o % 2) ? inv : -inv;
}

__device__ __forceinline__ void inv4(const double *m, double *out) {
  double inv[16];
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) inv[j * 4 + i] = cofactor4(i, j, m);
  double D = 0;
  for (int k = 0; k < 4; k++) D += m[k] * inv[k * 4];
  if (D == 0) return;
  const double Di = 1.0 / D;
  for (int i = 0; i < 16; i++) out[i] = inv[i] * Di;
}

__device__ __forceinline__ void matvec4(const double *A, const double *b, double *c) {
  for (int i = 0; i < 4; i++) {
    c[i] = 0.0;
    for (int j = 0; j < 4; j++)
      if (A[i * 4 + j] != 0.0 && b[j] != 0.0) c[i] += A[i * 4 + j] * b[j];
  }
}

__device__ __forceinline__ void gmres44(const double *A, const double *b, const double *x0, double *x) {
  constexpr int N = 4, MS = 4;
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
    const double root_inv =
        1.0 / sqrt(hm[j * MS + j] * hm[j * MS + j] + hm[(j + 1) * MS + j] * hm[(j + 1) * MS + j]);
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

__device__ __forceinline__ void cp3(double *d, const double *s) {
  d[0] = s[0];
  d[1] = s[1];
  d[2] = s[2];
}

__device__ __forceinline__ void cp4(double *d, const double *s) {
  d[0] = s[0];
  d[1] = s[1];
  d[2] = s[2];
  d[3] = s[3];
}

// calc_contact_point; returns Newton iterations, sets fail
__device__ __forceinline__ int contact_point(Sq &A, Sq &B, double ratio, const double *init1, double *result,
                                          double &fi, double &fj, int &fail, int max_iter) {
  const double tol1 = 1e-10, tol2 = 1e-12;
  fail = 0;
  double mu, mu_sq, F[4], F1[4], F2[4], point[3];
  double mu1, mu2, gA1[3], gB1[3], gA2[3], gB2[3];
  double fi1, fj1, fi2, fj2, merit01, merit02, merit0;
  const double res01 = calc_F_free(A, B, fi1, fj1, gA1, gB1, init1, mu1, F1, merit01);
  if (merit01 < tol1) {
    cp3(result, init1);
    cp3(A.gradient, gA1);
    cp3(B.gradient, gB1);
    fi = fi1;
    fj = fj1;
    return 0;
  }
  double init2[3];
  for (int k = 0; k < 3; k++) init2[k] = ratio * B.center[k] + (1.0 - ratio) * A.center[k];
  const double res02 = calc_F_free(A, B, fi2, fj2, gA2, gB2, init2, mu2, F2, merit02);
  double res0;
  if (res01 < res02) {
    res0 = res01;
    cp3(A.gradient, gA1);
    cp3(B.gradient, gB1);
    cp3(point, init1);
    cp4(F, F1);
    fi = fi1;
    fj = fj1;
    mu = mu1;
    merit0 = merit01;
  } else {
    res0 = res02;
    cp3(A.gradient, gA2);
    cp3(B.gradient, gB2);
    cp3(point, init2);
    cp4(F, F2);
    fi = fi2;
    fj = fj2;
    mu = mu2;
    merit0 = merit02;
  }
  if (merit0 < tol1) {
    cp3(result, point);
    return 0;
  }
  double merit1 = merit0, merit2 = merit0, res1 = res0, res2 = res1;
  const double size_i = fmin(fmin(A.shape[0], A.shape[1]), A.shape[2]);
  const double size_j = fmin(fmin(B.shape[0], B.shape[1]), B.shape[2]);
  const double size = size_i < size_j ? size_i : size_j;
  double J4[16], J4inv[16], delta[4] = {0, 0, 0, 0}, delta_0[4], pointa[3], pointb[3];
  int iter;
  for (iter = 0; iter < max_iter; iter++) {
    merit2 = merit1;
    res2 = res1;
    {
      double fd, gd[3];
      props_global<true>(A, point, fd, gd, A.hessian);
      props_global<true>(B, point, fd, gd, B.hessian);
    }
    mu_sq = mu * mu;
    for (int i = 0; i < 3; i++) {
      for (int j = 0; j < 3; j++) J4[4 * i + j] = A.hessian[i * 3 + j] + mu_sq * B.hessian[i * 3 + j];
      J4[3 + 4 * i] = 2.0 * mu * B.gradient[i];
      J4[i + 4 * 3] = A.gradient[i] - B.gradient[i];
    }
    J4[15] = 0.0;
    const double det = det4(J4);
    if (fabs(det) > 1.0) {
      inv4(J4, J4inv);
      matvec4(J4inv, F, delta);
    } else {
      cp4(delta_0, delta);
      gmres44(J4, F, delta_0, delta);
    }
    int nan_found = 0;
    for (int i = 0; i < 4; i++)
      if (isnan(delta[i])) {
        nan_found = 1;
        fail = 1;
        break;
      }
    int converged = 0;
    const double deltax = fmax(fmax(fabs(delta[0]), fabs(delta[1])), fabs(delta[2]));
    if (deltax == 0.0) {
      cp3(result, point);
      converged = 1;
    } else {
      if (nan_found) break;
      double point_[3], fi_, fj_, merit2_;
      for (int k = 0; k < 3; k++) point_[k] = point[k] - delta[k];
      const double mu_ = mu - delta[3];
      const double res2_ = calc_F_mu(A, B, fi_, fj_, A.gradient, B.gradient, point_, mu_, F, merit2_);
      if (res2_ < res1 || merit2_ < tol1 || deltax < tol2 * size) {
        merit2 = merit2_;
        res2 = res2_;
        mu = mu_;
        fi = fi_;
        fj = fj_;
        cp3(point, point_);
        if (merit2 < tol1 || deltax < tol2 * size) converged = 1;
      } else {
        double a = 0.0, b = 1.0, eps = fabs(b - a);
        double res2a = 0, res2b = 0, merit2a = 0, merit2b = 0;
        while (eps > 1e-8) {
          const double alpha1 = b - (b - a) * kPhiInv;
          const double alpha2 = a + (b - a) * kPhiInv;
          for (int k = 0; k < 3; k++) pointa[k] = point[k] + (-alpha1) * delta[k];
          mu1 = mu - alpha1 * delta[3];
          res2a = calc_F_mu(A, B, fi1, fj1, gA1, gB1, pointa, mu1, F1, merit2a);
          for (int k = 0; k < 3; k++) pointb[k] = point[k] + (-alpha2) * delta[k];
          mu2 = mu - alpha2 * delta[3];
          res2b = calc_F_mu(A, B, fi2, fj2, gA2, gB2, pointb, mu2, F2, merit2b);
          if ((res2a < res2b ? res2a : res2b) < res1) {
            if (res2a < res2b) {
              res2 = res2a;
              cp4(F, F1);
              mu = mu1;
              cp3(A.gradient, gA1);
              cp3(B.gradient, gB1);
              cp3(point, pointa);
              fi = fi1;
              fj = fj1;
              merit2 = merit2a;
            } else {
              res2 = res2b;
              cp4(F, F2);
              mu = mu2;
              cp3(A.gradient, gA2);
              cp3(B.gradient, gB2);
              cp3(point, pointb);
              fi = fi2;
              fj = fj2;
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
          cp4(F, F2);
          mu = mu2;
          cp3(A.gradient, gA2);
          cp3(B.gradient, gB2);
          cp3(point, pointb);
          fi = fi2;
          fj = fj2;
          res2 = res2b;
          merit2 = merit2b;
          if (merit2 < tol1) converged = 1;
        }
      }
      if (fabs(res1 - res2) / fmax(res1, res2) < 1e-3) converged = 1;
    }
    merit1 = merit2;
    res1 = res2;
    if (iter >= max_iter - 1) fail = 1;
    if (converged || fail) {
      cp3(result, point);
      break;
    }
  }
  return iter + 1;
}

// Extra full Newton steps on the converged system (F = 0, f_A = f_B) from
// point cp, each kept only if it lowers the residual; stops after `steps`,
// at a residual below 1e-30 or when a step does not help. Not part of the
// reference's solve: an optional refinement past its merit tolerance.
__device__ __forceinline__ int polish(Sq &A, Sq &B, double *cp, double &fi, double &fj, int steps) {
  double F[4], mu, merit, fA, fB, gA[3], gB[3];
  double res = calc_F_free(A, B, fA, fB, gA, gB, cp, mu, F, merit);
  cp3(A.gradient, gA);
  cp3(B.gradient, gB);
  int done = 0;
  for (int it = 0; it < steps; it++) {
    if (!(res > 1e-30)) break;
    double fd, gd[3], J4[16], J4inv[16], delta[4] = {0, 0, 0, 0}, d0[4] = {0, 0, 0, 0};
    props_global<true>(A, cp, fd, gd, A.hessian);
    props_global<true>(B, cp, fd, gd, B.hessian);
    const double mu_sq = mu * mu;
    for (int i = 0; i < 3; i++) {
      for (int j = 0; j < 3; j++) J4[4 * i + j] = A.hessian[i * 3 + j] + mu_sq * B.hessian[i * 3 + j];
      J4[3 + 4 * i] = 2.0 * mu * B.gradient[i];
      J4[i + 4 * 3] = A.gradient[i] - B.gradient[i];
    }
    J4[15] = 0.0;
    if (fabs(det4(J4)) > 1.0) {
      inv4(J4, J4inv);
      matvec4(J4inv, F, delta);
    } else {
      gmres44(J4, F, d0, delta);
    }
    double p_[3], F_[4], m_, fA_, fB_, gA_[3], gB_[3];
    for (int k = 0; k < 3; k++) p_[k] = cp[k] - delta[k];
    const double mu_ = mu - delta[3];
    const double r_ = calc_F_mu(A, B, fA_, fB_, gA_, gB_, p_, mu_, F_, m_);
    if (!(r_ < res)) break;
    res = r_;
    mu = mu_;
    cp3(cp, p_);
    cp4(F, F_);
    cp3(A.gradient, gA_);
    cp3(B.gradient, gB_);
    fi = fA_;
    fj = fB_;
    done++;
  }
  return done;
}

// the homotopy from volume-equivalent spheres (N+1 solves)
__device__ __forceinline__ int contact_point_fresh(Sq &A, Sq &B, double *cp, double &fi, double &fj, int &fail,
                                                int max_iter) {
  const double ai = A.shape[0], bi = A.shape[1], ci = A.shape[2];
  const double aj = B.shape[0], bj = B.shape[1], cj = B.shape[2];
  const double ri = cbrt(ai * bi * ci), rj = cbrt(aj * bj * cj);
  const double ratio = ri / (ri + rj);
  const double ni0 = A.blk[0], ni1 = A.blk[1], nj0 = B.blk[0], nj1 = B.blk[1];
  const double nmax = fmax(fmax(fmax(ni0, ni1), nj0), nj1);
  const int N1 = (int)((nmax - 2.0) / 1.0 + 1);
  const int N = N1 > 10 ? N1 : 10;
  const double Ninv = 1.0 / (double)N;
  int iters = 0, f = 0;
  fail = 0;
  for (int i = 0; i <= N; i++) {
    set_blockiness(A, 2.0 + (ni0 - 2.0) * (double)i * Ninv, 2.0 + (ni1 - 2.0) * (double)i * Ninv);
    set_blockiness(B, 2.0 + (nj0 - 2.0) * (double)i * Ninv, 2.0 + (nj1 - 2.0) * (double)i * Ninv);
    set_shape(A, ri + (ai - ri) * (double)i * Ninv, ri + (bi - ri) * (double)i * Ninv,
              ri + (ci - ri) * (double)i * Ninv);
    set_shape(B, rj + (aj - rj) * (double)i * Ninv, rj + (bj - rj) * (double)i * Ninv,
              rj + (cj - rj) * (double)i * Ninv);
    if (i == 0)
      for (int k = 0; k < 3; k++) cp[k] = ratio * B.center[k] + (1.0 - ratio) * A.center[k];
    const double pre[3] = {cp[0], cp[1], cp[2]};
    iters += contact_point(A, B, ratio, pre, cp, fi, fj, f, max_iter);
    fail = f;
    set_shape(A, ai, bi, ci);
    set_shape(B, aj, bj, cj);
    set_blockiness(A, ni0, ni1);
    set_blockiness(B, nj0, nj1);
  }
  return iters;
}

// ---------------------------------------------------------------------------
// contact detection around the contact point
// ---------------------------------------------------------------------------

// shape_function_local (superquadric.cpp:109-136): its own arithmetic, not
// props_local's
__device__ __forceinline__ double shape_local(const Sq &s, const double *x) {
  const double n1 = s.blk[0], n2 = s.blk[1];
  const double xi = x[0] * s.shape_inv[0], yi = x[1] * s.shape_inv[1], zi = x[2] * s.shape_inv[2];
  double f;
  if (s.int_blk) {
    const int n1i = (int)floor(n1 + 1e-2), n2i = (int)floor(n2 + 1e-2);
    if (n1i == n2i)
      f = pow_abs_int_sel(xi, n2i) + pow_abs_int_sel(yi, n2i) + pow_abs_int_sel(zi, n1i) - 1.0;
    else
      f = pow_abs(pow_abs_int_sel(xi, n2i) + pow_abs_int_sel(yi, n2i), n1 / n2) + pow_abs_int_sel(zi, n1i) - 1.0;
  } else {
    f = pow_abs(pow_abs(xi, n2) + pow_abs(yi, n2), n1 / n2) + pow_abs(zi, n1) - 1.0;
  }
  return s.koef * f;
}

// shape_function_gradient_local (superquadric.cpp:139-176)
__device__ __forceinline__ void gradient_local(const Sq &s, const double *x, double *r) {
  const double n1 = s.blk[0], n2 = s.blk[1];
  const double a = s.shape_inv[0], b = s.shape_inv[1], c = s.shape_inv[2];
  const double xa = x[0] * a, yb = x[1] * b, zc = x[2] * c;
  double xan21, ybn21, zcn11;
  if (s.int_blk) {
    const int n1i = (int)floor(n1 + 1e-2), n2i = (int)floor(n2 + 1e-2);
    xan21 = pow_abs_int_sel(xa, n2i - 1);
    ybn21 = pow_abs_int_sel(yb, n2i - 1);
    zcn11 = pow_abs_int_sel(zc, n1i - 1);
  } else {
    xan21 = pow_abs(xa, n2 - 1.0);
    ybn21 = pow_abs(yb, n2 - 1.0);
    zcn11 = pow_abs(zc, n1 - 1.0);
  }
  const double k = s.koef;
  double xyp1 = 1.0;
  const int same = comp_double(n1, n2, 1e-2);
  if (!same) xyp1 = pow_abs(xan21 * fabs(xa) + ybn21 * fabs(yb), n1 / n2 - 1.0);
  r[0] = same ? n1 * (k * a) * sgn(xa) * xan21 : n1 * (k * a) * sgn(xa) * xan21 * xyp1;
  r[1] = same ? n1 * (k * b) * sgn(yb) * ybn21 : n1 * (k * b) * sgn(yb) * ybn21 * xyp1;
  r[2] = n1 * (k * c) * sgn(zc) * zcn11;
}

__device__ __forceinline__ double clampd(double v, double lo, double hi) {
  return v < lo ? lo : (v > hi ? hi : v);
}

// Superquadric::surface_line_intersection (superquadric.cpp:712-816) with the
// ellipsoid's closed form (ellipsoid_line_intersection_local, :1011-1035),
// for the overlap algorithm's call (20 iterations, no ellipsoid start for
// other shapes): the alpha where start + alpha * dir crosses the surface
__device__ __forceinline__ double line_intersection(const Sq &s, const double *start, const double *dir,
                                                    double alpha_prev, double *result) {
  const double *R = s.rot;
  const double d0[3] = {start[0] - s.center[0], start[1] - s.center[1], start[2] - s.center[2]};
  const double sl[3] = {R[0] * d0[0] + R[3] * d0[1] + R[6] * d0[2], R[1] * d0[0] + R[4] * d0[1] + R[7] * d0[2],
                        R[2] * d0[0] + R[5] * d0[1] + R[8] * d0[2]};
  const double dl[3] = {R[0] * dir[0] + R[3] * dir[1] + R[6] * dir[2],
                        R[1] * dir[0] + R[4] * dir[1] + R[7] * dir[2],
                        R[2] * dir[0] + R[5] * dir[1] + R[8] * dir[2]};
  double alpha = alpha_prev;
  const double eps0 = 1e-14;
  const int ellipsoid = comp_double(s.blk[0], 2.0, 1e-2) && comp_double(s.blk[1], 2.0, 1e-2);
  if (ellipsoid) {
    const double x0 = sl[0] * s.shape_inv[0], nx = dl[0] * s.shape_inv[0];
    const double y0 = sl[1] * s.shape_inv[1], ny = dl[1] * s.shape_inv[1];
    const double z0 = sl[2] * s.shape_inv[2], nz = dl[2] * s.shape_inv[2];
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
    if (a1 > -eps0 && a2 < -eps0)
      alpha = a1;
    else if (a1 < -eps0 && a2 > -eps0)
      alpha = a2;
    else if (a1 < -eps0 && a2 < -eps0)
      alpha = a1 > a2 ? a1 : a2;
    else
      alpha = a1 < a2 ? a1 : a2;
  } else {
    const double a = 0.5 * fmin(fmin(s.shape[0], s.shape[1]), s.shape[2]) /
                     sqrt(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]);
    const double eps = 1e-16;
    alpha = clampd(alpha, -a, a);
    double point[3] = {sl[0] + alpha * dl[0], sl[1] + alpha * dl[1], sl[2] + alpha * dl[2]};
    double f0 = shape_local(s, point);
    if (!(fabs(f0) < eps)) {
      for (int i = 0; i < 20; i++) {
        if (fabs(f0) < eps * s.koef) break;
        double grad[3];
        gradient_local(s, point, grad);
        const double f_der = dot3(grad, dl);
        double delta;
        if (fabs(f_der) < 1e-10) {
          delta = 1e-10;
          const double p_[3] = {point[0] + delta * dl[0], point[1] + delta * dl[1], point[2] + delta * dl[2]};
          double g_[3];
          gradient_local(s, p_, g_);
          if (dot3(g_, dl) < 0.0) delta = -delta;
        } else {
          delta = clampd(-f0 / f_der, -a, a);
        }
        if (fabs(delta) <= eps) {
          alpha += delta;
          for (int k = 0; k < 3; k++) point[k] = point[k] + delta * dl[k];
          break;
        }
        while (fabs(delta) > eps) {
          const double p_[3] = {point[0] + delta * dl[0], point[1] + delta * dl[1], point[2] + delta * dl[2]};
          const double f_ = shape_local(s, p_);
          if (fabs(f_) < fabs(f0) || fabs(f_) < eps * s.koef) {
            f0 = f_;
            alpha += delta;
            for (int k = 0; k < 3; k++) point[k] = p_[k];
            break;
          }
          delta *= 0.5;
        }
      }
    } else {
      // already on the surface: the point mapped back (superquadric.cpp:766-771)
      for (int k = 0; k < 3; k++)
        result[k] = R[3 * k] * point[0] + R[3 * k + 1] * point[1] + R[3 * k + 2] * point[2] + s.center[k];
      return alpha;
    }
  }
  for (int k = 0; k < 3; k++) result[k] = start[k] + alpha * dir[k];
  return alpha;
}

// check_inequalities (math_extra_liggghts_superquadric.cpp:72-150): axis i
// separates the two boxes; C = R_A^T R_B, Ad = R_A^T d, Bd = R_B^T d
__device__ __forceinline__ int obb_separated(int i, const double *a, const double *b, const double *Ad,
                                             const double *Bd, const double *C) {
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
    default: R0 = a[0] * fabs(C[5]) + a[1] * fabs(C[2]); R1 = b[0] * fabs(C[7]) + b[1] * fabs(C[6]); R = fabs(C[2] * Ad[1] - C[5] * Ad[0]); break;
  }
  if (fabs(R0 + R1) < 1e-14 && fabs(R) < 1e-14) return 0;
  return R > R0 + R1;
}

// obb_intersect (math_extra_liggghts_superquadric.cpp:166-203): boxes of the
// half-axes grown by 1.0001 (scale_koef, math_extra_liggghts_nonspherical.h:60);
// the last separating axis is tried first
__device__ __forceinline__ int obb_intersect(const Sq &A, const Sq &B, int &start) {
  const double sk = 1.0001;
  const double d[3] = {A.center[0] - B.center[0], A.center[1] - B.center[1], A.center[2] - B.center[2]};
  const double a[3] = {A.shape[0] * sk, A.shape[1] * sk, A.shape[2] * sk};
  const double b[3] = {B.shape[0] * sk, B.shape[1] * sk, B.shape[2] * sk};
  const double *m = A.rot, *m2 = B.rot;
  double C[9];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) C[i * 3 + j] = m[i] * m2[j] + m[3 + i] * m2[3 + j] + m[6 + i] * m2[6 + j];
  const double Ad[3] = {m[0] * d[0] + m[3] * d[1] + m[6] * d[2], m[1] * d[0] + m[4] * d[1] + m[7] * d[2],
                        m[2] * d[0] + m[5] * d[1] + m[8] * d[2]};
  const double Bd[3] = {m2[0] * d[0] + m2[3] * d[1] + m2[6] * d[2], m2[1] * d[0] + m2[4] * d[1] + m2[7] * d[2],
                        m2[2] * d[0] + m2[5] * d[1] + m2[8] * d[2]};
  for (int i = 0; i < 15; i++) {
    const int which = i == 0 ? start : (i == start ? 0 : i);
    if (obb_separated(which, a, b, Ad, Bd, C)) {
      start = which;
      return 0;
    }
  }
  return 1;
}

}  // namespace sqd
