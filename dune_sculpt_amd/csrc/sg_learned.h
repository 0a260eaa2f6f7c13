// Learned Smooth model (host object + kernel view).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "sg_internal.h"
#include "sg_kernels.h"

#define SG_MLP_NPARAM 6211
#define SG_OFF_W1 0                      // [28][64]
#define SG_OFF_B1 (28 * 64)              // [64]
#define SG_OFF_W2 (SG_OFF_B1 + 64)       // [64][64]
#define SG_OFF_B2 (SG_OFF_W2 + 64 * 64)  // [64]
#define SG_OFF_W3 (SG_OFF_B2 + 64)       // [64][3]
#define SG_OFF_B3 (SG_OFF_W3 + 64 * 3)   // [3]

struct SgMlpView {
  const uint16_t *wb;  // bf16 W1t[64][32] | W2t[64][64] | W3t[16][64]
  const float *bias;   // b1[64] | b2[64] | b3[16]
  const float *norm;   // in_mu[28] | in_inv_sigma[28] | out_scale[3]
};

struct SgMlp {
  float *w = nullptr, *mo = nullptr, *ve = nullptr;  // FP32 master weights, Adam moments
  uint16_t *wb = nullptr;
  float *bias = nullptr, *norm = nullptr;
  float host_norm[64];
  float *grad_part = nullptr, *loss_part = nullptr;
  int part_cap = 0;
  float *dX = nullptr, *dT = nullptr;
  int *perm = nullptr;
  int n = 0;
  int step = 0;
  float alpha = 1.0f;  // held-out RMS error / RMS displacement of the plain brush
  float tol = 1e-3f;   // accepted error estimate, relative to the brush radius
  hipStream_t stream = nullptr;
};

SgMlpView sg_mlp_view(SgMlp *m);
int sg_enqueue_learned_brush(PBVH *p, const SgDevView &v, const SgDabK &k, const float *curve, SgMlp *m,
                             const float *place = nullptr);
int sg_enqueue_learned_post(PBVH *p, const SgDevView &v, SgMlp *m, int serial);
int sg_enqueue_collect(PBVH *p, const SgDevView &v, float *X, float *frame, int cap);
int sg_enqueue_judge(PBVH *p, int na, const float *lrn, const float *bx, const float *by, const float *bz,
                     double *acc);
int sg_enqueue_target(PBVH *p, const SgDevView &v, const float *frame, float *T, int n);
