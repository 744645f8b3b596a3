// Superquadric contact points on the GPU: the FP64 Newton solve for many
// independent pose pairs (one thread each), and the contact-point network
// that warm-starts it -- an 18 -> 64 -> 64 -> 3 tanh MLP whose forward and
// backward passes run on BF16 matrix cores (v_mfma_f32_16x16x16_bf16) with
// FP32 accumulation, trained here with Adam on FP32 master weights.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <mutex>

#include "sq_device.h"
#include "sq_kernels.h"

namespace sqk {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr int kBlk = 256;

// ---------------------------------------------------------------------------
// contact-point solve
// ---------------------------------------------------------------------------

// mode 0: fresh contacts (homotopy from spheres); mode 1: one solve from
// init[k], and the homotopy when that one fails or stops without meeting the
// merit tolerance (the plain start then judges and replaces it). The
// fallback stays in the same launch: run as a separate launch over the few
// listed pairs, the hardest homotopies would run alone on an idle GPU.
__global__ __launch_bounds__(kBlk) void k_sq_solve(SolveArgs a) {
  const int64_t k = (int64_t)blockIdx.x * kBlk + threadIdx.x;
  if (k >= a.n) return;
  sqd::Sq A, B;
  sqd::init(A, a.pose_a + 12 * k);
  sqd::init(B, a.pose_b + 12 * k);
  double cp[3], fi = 0, fj = 0;
  int fail = 0, iters = 0, fallback = 0;
  if (a.mode == 0) {
    iters = sqd::contact_point_fresh(A, B, cp, fi, fj, fail, a.max_iter);
  } else {
    const double ri = cbrt(A.shape[0] * A.shape[1] * A.shape[2]);
    const double rj = cbrt(B.shape[0] * B.shape[1] * B.shape[2]);
    const double init[3] = {a.init[3 * k], a.init[3 * k + 1], a.init[3 * k + 2]};
    iters = sqd::contact_point(A, B, ri / (ri + rj), init, cp, fi, fj, fail, a.max_iter);
    const double merit = sqd::merit_of(A.gradient, B.gradient, fi, fj);
    if (fail || !(merit < 1e-10)) {
      fallback = 1;
      iters += sqd::contact_point_fresh(A, B, cp, fi, fj, fail, a.max_iter);
    }
  }
  if (a.polish > 0 && !fail) iters += sqd::polish(A, B, cp, fi, fj, a.polish);
  a.point[3 * k] = cp[0];
  a.point[3 * k + 1] = cp[1];
  a.point[3 * k + 2] = cp[2];
  a.fi[k] = fi;
  a.fj[k] = fj;
  a.iters[k] = iters;
  a.fail[k] = fail | (fallback << 1);
}

// SurfaceModel<SURFACE_SUPERQUADRIC>::checkSurfaceIntersect
// (surface_model_superquadric.h:108-205): OBB pre-check unless the pair
// intersected last step, the contact point (homotopy for a pair far apart
// last step, else from the previous point moved by the mass-weighted contact
// velocity, calc_contact_point_using_prev_step, math_extra_liggghts_superquadric.cpp:
// 1434-1466), and for intersecting surfaces the normal from the two gradients
// and the overlap along it (extended_overlap_algorithm, :1489-1504). With
// init, a pair far apart last step starts from the network's point instead of
// the homotopy, which remains the fallback.
__global__ __launch_bounds__(kBlk) void k_sq_check(CheckArgs a) {
  const int64_t k = (int64_t)blockIdx.x * kBlk + threadIdx.x;
  if (k >= a.n) return;
  sqd::Sq A, B;
  sqd::init(A, a.pose_a + 12 * k);
  sqd::init(B, a.pose_b + 12 * k);
  double *h = a.hist + kHist * k;
  const int state = (int)(h[0] + 0.5);
  int start = (int)(h[1] + 0.5);
  const int obb = state == 2 ? 1 : sqd::obb_intersect(A, B, start);
  int contact = 0, iters = 0, fail = 0, fallback = 0;
  double cp[3] = {h[2], h[3], h[4]}, fi = 0.0, fj = 0.0;
  double en[3] = {0.0, 0.0, 0.0}, dn = 0.0, cpa[3] = {0.0, 0.0, 0.0}, cpb[3] = {0.0, 0.0, 0.0};
  if (obb) {
    const double ri = cbrt(A.shape[0] * A.shape[1] * A.shape[2]);
    const double rj = cbrt(B.shape[0] * B.shape[1] * B.shape[2]);
    const double ratio = ri / (ri + rj);
    int fresh = state == 0;
    if (!fresh || a.init) {
      double est[3];
      if (fresh) {
        for (int c = 0; c < 3; c++) est[c] = a.init[3 * k + c];
      } else if (a.vel && a.mass) {
        const double *v = a.vel + 12 * k;
        const double mi = a.mass[2 * k], mj = a.mass[2 * k + 1];
        const double xci[3] = {cp[0] - A.center[0], cp[1] - A.center[1], cp[2] - A.center[2]};
        const double xcj[3] = {cp[0] - B.center[0], cp[1] - B.center[1], cp[2] - B.center[2]};
        const double *wi = v + 6, *wj = v + 9;
        const double ri_[3] = {wi[1] * xci[2] - wi[2] * xci[1], wi[2] * xci[0] - wi[0] * xci[2],
                               wi[0] * xci[1] - wi[1] * xci[0]};
        const double rj_[3] = {wj[1] * xcj[2] - wj[2] * xcj[1], wj[2] * xcj[0] - wj[0] * xcj[2],
                               wj[0] * xcj[1] - wj[1] * xcj[0]};
        for (int c = 0; c < 3; c++) {
          const double vi = v[c] + ri_[c], vj = v[3 + c] + rj_[c];
          const double veff = (vi * mi + vj * mj) / (mi + mj);
          est[c] = cp[c] + veff * a.dt;
        }
      } else {
        for (int c = 0; c < 3; c++) est[c] = cp[c];
      }
      iters = sqd::contact_point(A, B, ratio, est, cp, fi, fj, fail, a.max_iter);
      if (fresh) {
        const double merit = sqd::merit_of(A.gradient, B.gradient, fi, fj);
        if (fail || !(merit < 1e-10)) fallback = 1;
        fresh = fallback;
      }
    }
    if (fresh) iters += sqd::contact_point_fresh(A, B, cp, fi, fj, fail, a.max_iter);
    contact = (fi > fj ? fi : fj) < 0.0;
    if (contact) {
      for (int c = 0; c < 3; c++) en[c] = B.gradient[c] - A.gradient[c];
      const double nrm = sqrt(en[0] * en[0] + en[1] * en[1] + en[2] * en[2]);
      const double inv = nrm == 0. ? 0. : 1. / nrm;
      for (int c = 0; c < 3; c++) en[c] *= inv;
      const double men[3] = {-en[0], -en[1], -en[2]};
      h[5] = sqd::line_intersection(A, cp, men, h[5], cpa);
      h[6] = sqd::line_intersection(B, cp, en, h[6], cpb);
      dn = fabs(h[5]) + fabs(h[6]);
      h[0] = 2.0;
    } else {
      h[0] = 1.0;
    }
    h[2] = cp[0];
    h[3] = cp[1];
    h[4] = cp[2];
  } else {
    h[0] = 0.0;
  }
  h[1] = (double)start;
  a.contact[k] = contact;
  for (int c = 0; c < 3; c++) {
    a.en[3 * k + c] = en[c];
    a.cp_i[3 * k + c] = cpa[c];
    a.cp_j[3 * k + c] = cpb[c];
    a.point[3 * k + c] = cp[c];
  }
  a.deltan[k] = dn;
  a.fi[k] = fi;
  a.fj[k] = fj;
  a.iters[k] = iters;
  a.flags[k] = fail | (fallback << 1) | ((!obb) << 2);
}

// ---------------------------------------------------------------------------
// network inputs and outputs (the encoding of superquadrics.py)
// ---------------------------------------------------------------------------
__device__ __forceinline__ void rot_of(const double *q, double *R) {
  const double w = q[0], i = q[1], j = q[2], kk = q[3];
  R[0] = w * w + i * i - j * j - kk * kk;
  R[1] = 2 * i * j - 2 * kk * w;
  R[2] = 2 * j * w + 2 * i * kk;
  R[3] = 2 * i * j + 2 * kk * w;
  R[4] = w * w - i * i + j * j - kk * kk;
  R[5] = 2 * j * kk - 2 * i * w;
  R[6] = 2 * i * kk - 2 * j * w;
  R[7] = 2 * j * kk + 2 * i * w;
  R[8] = w * w - i * i - j * j + kk * kk;
}

__global__ __launch_bounds__(kBlk) void k_sq_features(const double *pa, const double *pb, int64_t n,
                                                       float *X) {
  const int64_t k = (int64_t)blockIdx.x * kBlk + threadIdx.x;
  if (k >= n) return;
  const double *A = pa + 12 * k, *B = pb + 12 * k;
  const double rA = cbrt(A[7] * A[8] * A[9]), rB = cbrt(B[7] * B[8] * B[9]);
  const double L = rA + rB;
  double R[9];
  rot_of(A + 3, R);
  const double d[3] = {B[0] - A[0], B[1] - A[1], B[2] - A[2]};
  float *x = X + 18 * k;
  for (int c = 0; c < 3; c++) x[c] = (float)((R[c] * d[0] + R[3 + c] * d[1] + R[6 + c] * d[2]) / L);
  // conj(qA) * qB
  const double w1 = A[3], x1 = -A[4], y1 = -A[5], z1 = -A[6];
  const double w2 = B[3], x2 = B[4], y2 = B[5], z2 = B[6];
  double qr[4] = {w1 * w2 - x1 * x2 - y1 * y2 - z1 * z2, w1 * x2 + x1 * w2 + y1 * z2 - z1 * y2,
                  w1 * y2 - x1 * z2 + y1 * w2 + z1 * x2, w1 * z2 + x1 * y2 - y1 * x2 + z1 * w2};
  const double sg = qr[0] < 0.0 ? -1.0 : 1.0;
  for (int c = 0; c < 4; c++) x[3 + c] = (float)(qr[c] * sg);
  for (int c = 0; c < 3; c++) x[7 + c] = (float)(A[7 + c] / rA);
  x[10] = (float)(A[10] / 8.0);
  x[11] = (float)(A[11] / 8.0);
  for (int c = 0; c < 3; c++) x[12 + c] = (float)(B[7 + c] / rB);
  x[15] = (float)(B[10] / 8.0);
  x[16] = (float)(B[11] / 8.0);
  x[17] = (float)(rA / (rA + rB));
}

// global contact points from outputs (to_point), or targets from points (from_point)
__global__ __launch_bounds__(kBlk) void k_sq_frame(const double *pa, const double *pb, int64_t n,
                                                    const float *Y, double *P, const double *Pin,
                                                    float *T) {
  const int64_t k = (int64_t)blockIdx.x * kBlk + threadIdx.x;
  if (k >= n) return;
  const double *A = pa + 12 * k, *B = pb + 12 * k;
  const double L = cbrt(A[7] * A[8] * A[9]) + cbrt(B[7] * B[8] * B[9]);
  double R[9];
  rot_of(A + 3, R);
  if (Y) {
    const double y[3] = {(double)Y[3 * k] * L, (double)Y[3 * k + 1] * L, (double)Y[3 * k + 2] * L};
    for (int c = 0; c < 3; c++) P[3 * k + c] = A[c] + (R[3 * c] * y[0] + R[3 * c + 1] * y[1] + R[3 * c + 2] * y[2]);
  } else {
    const double d[3] = {Pin[3 * k] - A[0], Pin[3 * k + 1] - A[1], Pin[3 * k + 2] - A[2]};
    for (int c = 0; c < 3; c++) T[3 * k + c] = (float)((R[c] * d[0] + R[3 + c] * d[1] + R[6 + c] * d[2]) / L);
  }
}

// ---------------------------------------------------------------------------
// the network on matrix cores
//
// v_mfma_f32_16x16x16_bf16 lane layout (64 lanes, l = lane, r = l % 16,
// q = l / 16): A (16 x 16, M x K) element A[r][4q + j] in slot j of the lane's
// bf16x4; B (K x N) element B[4q + j][r]; the FP32 result D[4q + j][r] in
// slot j. One wavefront owns a tile of 16 samples.
// ---------------------------------------------------------------------------
__device__ __forceinline__ unsigned short to_bf16(float f) { return pack_bf16(f); }

__device__ __forceinline__ float from_bf16(unsigned short h) { return __uint_as_float((unsigned int)h << 16); }

__device__ __forceinline__ floatx4 mfma(bf16x4 a, bf16x4 b, floatx4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// A operand: rows r0.., columns k0.. of a row-major bf16 matrix (leading dim ld)
__device__ __forceinline__ bf16x4 ld_a(const unsigned short *M, int ld, int r0, int k0) {
  const int l = threadIdx.x & 63;
  const unsigned short *p = M + (r0 + (l & 15)) * ld + k0 + 4 * (l >> 4);
  bf16x4 v;
  v[0] = (short)p[0];
  v[1] = (short)p[1];
  v[2] = (short)p[2];
  v[3] = (short)p[3];
  return v;
}

// A operand of a transposed matrix: A[m][k] = M[k][m]
__device__ __forceinline__ bf16x4 ld_at(const unsigned short *M, int ld, int m0, int k0) {
  const int l = threadIdx.x & 63;
  const int m = m0 + (l & 15), k = k0 + 4 * (l >> 4);
  bf16x4 v;
  v[0] = (short)M[(k + 0) * ld + m];
  v[1] = (short)M[(k + 1) * ld + m];
  v[2] = (short)M[(k + 2) * ld + m];
  v[3] = (short)M[(k + 3) * ld + m];
  return v;
}

// B operand: B[k][n] of a row-major bf16 matrix
__device__ __forceinline__ bf16x4 ld_b(const unsigned short *M, int ld, int k0, int n0) {
  const int l = threadIdx.x & 63;
  const int n = n0 + (l & 15), k = k0 + 4 * (l >> 4);
  bf16x4 v;
  v[0] = (short)M[(k + 0) * ld + n];
  v[1] = (short)M[(k + 1) * ld + n];
  v[2] = (short)M[(k + 2) * ld + n];
  v[3] = (short)M[(k + 3) * ld + n];
  return v;
}

// B operand of a transposed matrix: B[k][n] = M[n][k]
__device__ __forceinline__ bf16x4 ld_bt(const unsigned short *M, int ld, int k0, int n0) {
  const int l = threadIdx.x & 63;
  const unsigned short *p = M + (n0 + (l & 15)) * ld + k0 + 4 * (l >> 4);
  bf16x4 v;
  v[0] = (short)p[0];
  v[1] = (short)p[1];
  v[2] = (short)p[2];
  v[3] = (short)p[3];
  return v;
}

struct WaveLds {
  unsigned short xb[16 * kInPad];
  unsigned short h1b[16 * kHid], h2b[16 * kHid], dzb[16 * kHid], dyb[16 * kOutPad];
};

// FP32 hidden activations of a tile in the result layout: h[nt][j] is
// row 4q + j, column 16 nt + r of the layer
struct Hidden {
  floatx4 h1[kHid / 16], h2[kHid / 16];
};

struct BlockLds {
  unsigned short w1[kInPad * kHid], w2[kHid * kHid], w3[kHid * kOutPad];
  float b1[kHid], b2[kHid], b3[kOutPad];
  WaveLds wave[kBlk / 64];
};

__device__ void stage_weights(BlockLds &s, const NetDev &net) {
  for (int i = threadIdx.x; i < kInPad * kHid; i += kBlk) s.w1[i] = net.w1b[i];
  for (int i = threadIdx.x; i < kHid * kHid; i += kBlk) s.w2[i] = net.w2b[i];
  for (int i = threadIdx.x; i < kHid * kOutPad; i += kBlk) s.w3[i] = net.w3b[i];
  for (int i = threadIdx.x; i < kHid; i += kBlk) {
    s.b1[i] = net.b1[i];
    s.b2[i] = net.b2[i];
  }
  for (int i = threadIdx.x; i < kOutPad; i += kBlk) s.b3[i] = net.b3[i];
}

// forward pass of one 16-sample tile; rows >= nvalid are zero inputs.
// Leaves the activations in w; returns this lane's outputs D[4q + j][r].
__device__ floatx4 forward_tile(const BlockLds &s, WaveLds &w, const float *X, const int64_t *rows,
                                int nvalid, Hidden &hid) {
  const int l = threadIdx.x & 63, r = l & 15, q = l >> 4;
  // inputs -> bf16, zero padded to kInPad columns
  for (int e = l; e < 16 * kInPad; e += 64) {
    const int row = e / kInPad, col = e % kInPad;
    float v = 0.f;
    if (row < nvalid && col < kIn) v = X[rows[row] * kIn + col];
    w.xb[e] = to_bf16(v);
  }
  __syncthreads();
  // layer 1
  for (int nt = 0; nt < kHid / 16; nt++) {
    floatx4 acc = {0.f, 0.f, 0.f, 0.f};
    for (int ks = 0; ks < kInPad / 16; ks++) acc = mfma(ld_a(w.xb, kInPad, 0, ks * 16), ld_b(s.w1, kHid, ks * 16, nt * 16), acc);
    for (int j = 0; j < 4; j++) {
      const int row = 4 * q + j, col = nt * 16 + r;
      const float h = tanhf(acc[j] + s.b1[col]);
      hid.h1[nt][j] = h;
      w.h1b[row * kHid + col] = to_bf16(h);
    }
  }
  __syncthreads();
  // layer 2
  for (int nt = 0; nt < kHid / 16; nt++) {
    floatx4 acc = {0.f, 0.f, 0.f, 0.f};
    for (int ks = 0; ks < kHid / 16; ks++) acc = mfma(ld_a(w.h1b, kHid, 0, ks * 16), ld_b(s.w2, kHid, ks * 16, nt * 16), acc);
    for (int j = 0; j < 4; j++) {
      const int row = 4 * q + j, col = nt * 16 + r;
      const float h = tanhf(acc[j] + s.b2[col]);
      hid.h2[nt][j] = h;
      w.h2b[row * kHid + col] = to_bf16(h);
    }
  }
  __syncthreads();
  // layer 3 (linear)
  floatx4 y = {0.f, 0.f, 0.f, 0.f};
  for (int ks = 0; ks < kHid / 16; ks++) y = mfma(ld_a(w.h2b, kHid, 0, ks * 16), ld_b(s.w3, kOutPad, ks * 16, 0), y);
  for (int j = 0; j < 4; j++) y[j] += s.b3[r];
  return y;
}

__global__ __launch_bounds__(kBlk) void k_mlp_forward(NetDev net, const float *X, int64_t n, float *Y) {
  __shared__ BlockLds s;
  stage_weights(s, net);
  __syncthreads();
  const int wv = threadIdx.x >> 6, l = threadIdx.x & 63, r = l & 15, q = l >> 4;
  const int64_t t0 = ((int64_t)blockIdx.x * (kBlk / 64) + wv) * 16;
  int64_t rows[16];
  for (int i = 0; i < 16; i++) rows[i] = t0 + i;
  const int nvalid = (int)(n - t0 < 0 ? 0 : (n - t0 > 16 ? 16 : n - t0));
  Hidden hid;
  const floatx4 y = forward_tile(s, s.wave[wv], X, rows, nvalid, hid);
  if (r < kOut)
    for (int j = 0; j < 4; j++) {
      const int row = 4 * q + j;
      if (row < nvalid) Y[(t0 + row) * kOut + r] = y[j];
    }
}

// one gradient pass: every wavefront takes one 16-sample tile of the batch
// (sample k of the batch is rows idx[k]), runs forward and backward with the
// products on matrix cores, and the block writes the sum of its tiles'
// parameter gradients to part[block][] (wavefronts added in a fixed order, so
// the result does not depend on timing). Loss = mean over the batch of
// |y - t|^2; BF16 copies feed every product: inputs, weights, hidden
// activations, and the output and hidden-layer deltas.
__device__ __forceinline__ float sum_over_q(float v) {
  v += __shfl_xor(v, 16, 64);
  v += __shfl_xor(v, 32, 64);
  return v;
}

__global__ __launch_bounds__(kBlk) void k_mlp_grad(NetDev net, const float *X, const float *T,
                                                    const int64_t *idx, int64_t batch, float *part,
                                                    float *loss_part) {
  __shared__ BlockLds s;
  __shared__ float red[kNumParams];
  __shared__ float lred[kBlk / 64];
  stage_weights(s, net);
  for (int i = threadIdx.x; i < kNumParams; i += kBlk) red[i] = 0.f;
  __syncthreads();
  const int wv = threadIdx.x >> 6, l = threadIdx.x & 63, r = l & 15, q = l >> 4;
  WaveLds &w = s.wave[wv];
  const int64_t t0 = ((int64_t)blockIdx.x * (kBlk / 64) + wv) * 16;
  const int nvalid = (int)(batch - t0 < 0 ? 0 : (batch - t0 > 16 ? 16 : batch - t0));
  int64_t rows[16];
  for (int i = 0; i < 16; i++) rows[i] = i < nvalid ? (idx ? idx[t0 + i] : t0 + i) : 0;
  Hidden hid;
  const floatx4 y = forward_tile(s, w, X, rows, nvalid, hid);
  const float scale = 2.0f / (float)batch;
  float lsum = 0.f, db3 = 0.f;
  for (int j = 0; j < 4; j++) {
    const int row = 4 * q + j;
    float d = 0.f;
    if (row < nvalid && r < kOut) {
      const float e = y[j] - T[rows[row] * kOut + r];
      lsum += e * e;
      d = scale * e;
    }
    db3 += d;
    w.dyb[row * kOutPad + r] = to_bf16(d);
  }
  __syncthreads();
  // dW3 = H2^T dY (kHid x kOut)
  floatx4 gw3[kHid / 16];
  for (int mt = 0; mt < kHid / 16; mt++) {
    const floatx4 z = {0.f, 0.f, 0.f, 0.f};
    gw3[mt] = mfma(ld_at(w.h2b, kHid, mt * 16, 0), ld_b(w.dyb, kOutPad, 0, 0), z);
  }
  // dZ2 = (dY W3^T) * (1 - h2^2)
  float db2[kHid / 16];
  for (int nt = 0; nt < kHid / 16; nt++) {
    floatx4 acc = {0.f, 0.f, 0.f, 0.f};
    acc = mfma(ld_a(w.dyb, kOutPad, 0, 0), ld_bt(s.w3, kOutPad, 0, nt * 16), acc);
    db2[nt] = 0.f;
    for (int j = 0; j < 4; j++) {
      const int row = 4 * q + j, col = nt * 16 + r;
      const float h = hid.h2[nt][j];
      const float dz = acc[j] * (1.0f - h * h);
      db2[nt] += dz;
      w.dzb[row * kHid + col] = to_bf16(dz);
    }
  }
  __syncthreads();
  // dW2 = H1^T dZ2 (kHid x kHid)
  floatx4 gw2[kHid / 16][kHid / 16];
  for (int mt = 0; mt < kHid / 16; mt++)
    for (int nt = 0; nt < kHid / 16; nt++) {
      const floatx4 z = {0.f, 0.f, 0.f, 0.f};
      gw2[mt][nt] = mfma(ld_at(w.h1b, kHid, mt * 16, 0), ld_b(w.dzb, kHid, 0, nt * 16), z);
    }
  // dH1 = dZ2 W2^T
  floatx4 dh1[kHid / 16];
  for (int nt = 0; nt < kHid / 16; nt++) {
    floatx4 acc = {0.f, 0.f, 0.f, 0.f};
    for (int ks = 0; ks < kHid / 16; ks++)
      acc = mfma(ld_a(w.dzb, kHid, 0, ks * 16), ld_bt(s.w2, kHid, ks * 16, nt * 16), acc);
    dh1[nt] = acc;
  }
  __syncthreads();  // every read of dZ2 is done
  float db1[kHid / 16];
  for (int nt = 0; nt < kHid / 16; nt++) {
    db1[nt] = 0.f;
    for (int j = 0; j < 4; j++) {
      const int row = 4 * q + j, col = nt * 16 + r;
      const float h = hid.h1[nt][j];
      const float dz = dh1[nt][j] * (1.0f - h * h);
      db1[nt] += dz;
      w.dzb[row * kHid + col] = to_bf16(dz);
    }
  }
  __syncthreads();
  // dW1 = X^T dZ1 (kInPad x kHid; rows >= kIn are padding)
  floatx4 gw1[kInPad / 16][kHid / 16];
  for (int mt = 0; mt < kInPad / 16; mt++)
    for (int nt = 0; nt < kHid / 16; nt++) {
      const floatx4 z = {0.f, 0.f, 0.f, 0.f};
      gw1[mt][nt] = mfma(ld_at(w.xb, kInPad, mt * 16, 0), ld_b(w.dzb, kHid, 0, nt * 16), z);
    }
  // bias sums over the tile's rows (lanes r, r+16, r+32, r+48)
  db3 = sum_over_q(db3);
  for (int nt = 0; nt < kHid / 16; nt++) {
    db2[nt] = sum_over_q(db2[nt]);
    db1[nt] = sum_over_q(db1[nt]);
  }
  for (int off = 32; off > 0; off >>= 1) lsum += __shfl_xor(lsum, off, 64);
  // the wavefronts add into the block sum one after the other
  for (int p = 0; p < kBlk / 64; p++) {
    if (wv == p) {
      for (int mt = 0; mt < kHid / 16; mt++)
        if (r < kOut)
          for (int j = 0; j < 4; j++) red[kOffW3 + (mt * 16 + 4 * q + j) * kOut + r] += gw3[mt][j];
      for (int mt = 0; mt < kHid / 16; mt++)
        for (int nt = 0; nt < kHid / 16; nt++)
          for (int j = 0; j < 4; j++) red[kOffW2 + (mt * 16 + 4 * q + j) * kHid + nt * 16 + r] += gw2[mt][nt][j];
      for (int mt = 0; mt < kInPad / 16; mt++)
        for (int nt = 0; nt < kHid / 16; nt++)
          for (int j = 0; j < 4; j++) {
            const int m = mt * 16 + 4 * q + j;
            if (m < kIn) red[kOffW1 + m * kHid + nt * 16 + r] += gw1[mt][nt][j];
          }
      if (q == 0) {
        if (r < kOut) red[kOffB3 + r] += db3;
        for (int nt = 0; nt < kHid / 16; nt++) {
          red[kOffB2 + nt * 16 + r] += db2[nt];
          red[kOffB1 + nt * 16 + r] += db1[nt];
        }
      }
      if (l == 0) lred[wv] = lsum;
    }
    __syncthreads();
  }
  for (int i = threadIdx.x; i < kNumParams; i += kBlk) part[(int64_t)blockIdx.x * kNumParams + i] = red[i];
  if (threadIdx.x == 0) {
    float v = 0.f;
    for (int k = 0; k < kBlk / 64; k++) v += lred[k];
    loss_part[blockIdx.x] = v;
  }
}

// sums the blocks' gradient partials (fixed order: deterministic), then one
// Adam step on the FP32 master parameters and a refresh of the BF16 / padded
// copies the matrix cores read. adam = 0: only the summed gradient is written.
__global__ __launch_bounds__(kBlk) void k_mlp_adam(NetDev net, const float *part, int nparts, float *grad,
                                                   AdamArgs ad) {
  const int i = blockIdx.x * kBlk + threadIdx.x;
  if (i == 0 && ad.loss_out) {  // the batch loss of this step
    float v = 0.f;
    for (int b = 0; b < nparts; b++) v += ad.loss_part[b];
    *ad.loss_out = v / (float)ad.batch;
  }
  if (i >= kNumParams) return;
  float gsum = 0.f;
  for (int b = 0; b < nparts; b++) gsum += part[(int64_t)b * kNumParams + i];
  grad[i] = gsum;
  if (!ad.apply) return;
  float m = ad.beta1 * net.m[i] + (1.0f - ad.beta1) * gsum;
  float v = ad.beta2 * net.v[i] + (1.0f - ad.beta2) * gsum * gsum;
  net.m[i] = m;
  net.v[i] = v;
  const float mh = m / ad.bc1, vh = v / ad.bc2;
  const float p = net.params[i] - ad.lr * mh / (sqrtf(vh) + ad.eps);
  net.params[i] = p;
  write_packed(net, i, p);
}

__global__ __launch_bounds__(kBlk) void k_mlp_pack(NetDev net) {
  const int i = blockIdx.x * kBlk + threadIdx.x;
  if (i < kNumParams) write_packed(net, i, net.params[i]);
}

// batch rows for a training step: uniform with replacement from [0, n)
__global__ __launch_bounds__(kBlk) void k_sample(int64_t *idx, int64_t batch, int64_t n, uint64_t seed) {
  const int64_t k = (int64_t)blockIdx.x * kBlk + threadIdx.x;
  if (k >= batch) return;
  uint64_t z = seed + 0x9e3779b97f4a7c15ull * (uint64_t)(k + 1);  // splitmix64
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  z ^= z >> 31;
  idx[k] = (int64_t)(z % (uint64_t)n);
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static inline unsigned nblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

hipError_t launch_solve(const SolveArgs &a, hipStream_t s) {
  if (a.n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_sq_solve, dim3(nblk(a.n, kBlk)), dim3(kBlk), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_check(const CheckArgs &a, hipStream_t s) {
  if (a.n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_sq_check, dim3(nblk(a.n, kBlk)), dim3(kBlk), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_features(const double *pa, const double *pb, int64_t n, float *X, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_sq_features, dim3(nblk(n, kBlk)), dim3(kBlk), 0, s, pa, pb, n, X);
  return hipGetLastError();
}

hipError_t launch_to_point(const double *pa, const double *pb, int64_t n, const float *Y, double *P,
                           hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_sq_frame, dim3(nblk(n, kBlk)), dim3(kBlk), 0, s, pa, pb, n, Y, P,
                     (const double *)nullptr, (float *)nullptr);
  return hipGetLastError();
}

hipError_t launch_to_target(const double *pa, const double *pb, int64_t n, const double *P, float *T,
                            hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_sq_frame, dim3(nblk(n, kBlk)), dim3(kBlk), 0, s, pa, pb, n,
                     (const float *)nullptr, (double *)nullptr, P, T);
  return hipGetLastError();
}

hipError_t launch_mlp_forward(const NetDev &net, const float *X, int64_t n, float *Y, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_mlp_forward, dim3(nblk(n, kBlk / 4)), dim3(kBlk), 0, s, net, X, n, Y);
  return hipGetLastError();
}

int grad_blocks(int64_t batch) { return (int)nblk(batch, kBlk / 4); }

// Workspace of the allocation-free gradient entry point: k_mlp_grad's
// per-block gradient partials ([blocks][kNumParams]) followed by its per-block
// loss partials, for batches of up to kGradScratchMaxBatch samples. Every
// block writes its whole partial, so nothing needs clearing between calls.
// Each device holds its own copy until the process ends.
constexpr int kGradScratchBlocks = (int)((kGradScratchMaxBatch + kBlk / 4 - 1) / (kBlk / 4));
static __device__ __align__(256) float scratch[(size_t)kGradScratchBlocks * (kNumParams + 1)];

// Device address of this device's scratch copy (the array's name is only a
// host-side stand-in). Resolved once per device with that device current.
hipError_t grad_scratch(int device, float **part, float **loss_part) {
  static std::mutex mu;
  static float *addr[kMaxScratchDevices] = {};
  if (device < 0 || device >= kMaxScratchDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(mu);
  if (!addr[device]) {
    float *p = nullptr;
    const hipError_t e = hipGetSymbolAddress((void **)&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    addr[device] = p;
  }
  *part = addr[device];
  *loss_part = addr[device] + (size_t)kGradScratchBlocks * kNumParams;
  return hipSuccess;
}

hipError_t launch_mlp_grad(const NetDev &net, const float *X, const float *T, const int64_t *idx,
                           int64_t batch, float *part, float *loss_part, hipStream_t s) {
  if (batch <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_mlp_grad, dim3(grad_blocks(batch)), dim3(kBlk), 0, s, net, X, T, idx, batch, part,
                     loss_part);
  return hipGetLastError();
}

hipError_t launch_mlp_adam(const NetDev &net, const float *part, int nparts, float *grad, const AdamArgs &ad,
                           hipStream_t s) {
  hipLaunchKernelGGL(k_mlp_adam, dim3(nblk(kNumParams, kBlk)), dim3(kBlk), 0, s, net, part, nparts, grad, ad);
  return hipGetLastError();
}

hipError_t launch_sample(int64_t *idx, int64_t batch, int64_t n, uint64_t seed, hipStream_t s) {
  if (batch <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_sample, dim3(nblk(batch, kBlk)), dim3(kBlk), 0, s, idx, batch, n, seed);
  return hipGetLastError();
}

hipError_t launch_mlp_pack(const NetDev &net, hipStream_t s) {
  hipLaunchKernelGGL(k_mlp_pack, dim3(nblk(kNumParams, kBlk)), dim3(kBlk), 0, s, net);
  return hipGetLastError();
}

}  // namespace sqk
