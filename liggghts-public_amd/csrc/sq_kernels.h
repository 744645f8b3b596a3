// Superquadric contact-point solve and the warm-start network: device data
// structures and launchers (sq_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sqk {

// network shape: kIn -> kHid -> kHid -> kOut, tanh hidden layers
constexpr int kIn = 18, kHid = 64, kOut = 3;
constexpr int kInPad = 32, kOutPad = 16;  // K and N padded to the 16-wide matrix-core tiles
// flat FP32 parameter vector: W1 [kIn][kHid], b1, W2 [kHid][kHid], b2, W3 [kHid][kOut], b3
constexpr int kOffW1 = 0;
constexpr int kOffB1 = kOffW1 + kIn * kHid;
constexpr int kOffW2 = kOffB1 + kHid;
constexpr int kOffB2 = kOffW2 + kHid * kHid;
constexpr int kOffW3 = kOffB2 + kHid;
constexpr int kOffB3 = kOffW3 + kHid * kOut;
constexpr int kNumParams = kOffB3 + kOut;

struct NetDev {
  float *params, *m, *v;            // FP32 master parameters and Adam moments
  unsigned short *w1b, *w2b, *w3b;  // BF16 weights, [kInPad][kHid], [kHid][kHid], [kHid][kOutPad]
  float *b1, *b2, *b3;              // FP32 biases, b3 padded to kOutPad
};

struct AdamArgs {
  int apply;
  float lr, beta1, beta2, eps, bc1, bc2;  // bc = 1 - beta^t
  const float *loss_part;                 // per-block loss sums of the gradient pass
  float *loss_out;                        // the step's mean loss, or null
  int64_t batch;
};

struct SolveArgs {
  const double *pose_a, *pose_b;  // [n][12]
  const double *init;             // [n][3], mode 1
  int64_t n;
  int mode;                       // 0 fresh (homotopy), 1 from init with the plain fallback
  int max_iter;
  int polish;                     // extra Newton refinement steps after the solve (0: none)
  double *point, *fi, *fj;
  int *iters, *fail;              // fail: bit 0 the solve's flag, bit 1 the fallback ran
};

// checkSurfaceIntersect over pose pairs; hist [n][7] = state (0 far, 1 close,
// 2 intersecting), first OBB inequality, contact point[3], alpha_i, alpha_j
constexpr int kHist = 7;
struct CheckArgs {
  const double *pose_a, *pose_b;  // [n][12]
  const double *init;             // [n][3] network start points for pairs far apart last step, or null
  const double *vel, *mass;       // [n][12] v_i, v_j, omega_i, omega_j and [n][2], or null
  double dt;
  int64_t n;
  int max_iter;
  double *hist;
  int *contact;
  double *en, *deltan, *cp_i, *cp_j;  // [n][3], [n], [n][3], [n][3]
  double *point, *fi, *fj;
  int *iters, *flags;  // flags: bit 0 the solve's flag, bit 1 the fallback ran, bit 2 no solve (OBB apart)
};

__device__ __forceinline__ unsigned short pack_bf16(float f) {
  unsigned int u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u) return (unsigned short)((u >> 16) | ((u & 0xffff) ? 0x40 : 0));
  u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
  return (unsigned short)(u >> 16);
}

// parameter i -> the copy the matrix cores read
__device__ __forceinline__ void write_packed(const NetDev &net, int i, float p) {
  if (i < kOffB1) {
    net.w1b[(i / kHid) * kHid + i % kHid] = pack_bf16(p);
  } else if (i < kOffW2) {
    net.b1[i - kOffB1] = p;
  } else if (i < kOffB2) {
    net.w2b[i - kOffW2] = pack_bf16(p);
  } else if (i < kOffW3) {
    net.b2[i - kOffB2] = p;
  } else if (i < kOffB3) {
    const int k = i - kOffW3;
    net.w3b[(k / kOut) * kOutPad + k % kOut] = pack_bf16(p);
  } else {
    net.b3[i - kOffB3] = p;
  }
}

hipError_t launch_solve(const SolveArgs &a, hipStream_t s);
hipError_t launch_check(const CheckArgs &a, hipStream_t s);
hipError_t launch_features(const double *pa, const double *pb, int64_t n, float *X, hipStream_t s);
hipError_t launch_to_point(const double *pa, const double *pb, int64_t n, const float *Y, double *P,
                           hipStream_t s);
hipError_t launch_to_target(const double *pa, const double *pb, int64_t n, const double *P, float *T,
                            hipStream_t s);
hipError_t launch_mlp_forward(const NetDev &net, const float *X, int64_t n, float *Y, hipStream_t s);
int grad_blocks(int64_t batch);
// idx: the batch's rows of X / T, or null for rows 0 .. batch-1
hipError_t launch_mlp_grad(const NetDev &net, const float *X, const float *T, const int64_t *idx,
                           int64_t batch, float *part, float *loss_part, hipStream_t s);
hipError_t launch_mlp_adam(const NetDev &net, const float *part, int nparts, float *grad, const AdamArgs &ad,
                           hipStream_t s);
hipError_t launch_mlp_pack(const NetDev &net, hipStream_t s);
hipError_t launch_sample(int64_t *idx, int64_t batch, int64_t n, uint64_t seed, hipStream_t s);

// the allocation-free gradient's workspace (a file-scope device array in
// sq_kernels.hip): batches up to kGradScratchMaxBatch samples, devices
// 0 .. kMaxScratchDevices-1; resolve with the device current
constexpr int64_t kGradScratchMaxBatch = 16384;
constexpr int kMaxScratchDevices = 64;
hipError_t grad_scratch(int device, float **part, float **loss_part);

}  // namespace sqk
