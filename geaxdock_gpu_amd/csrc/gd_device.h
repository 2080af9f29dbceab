// gd_device.h -- device-side building blocks of the MI355X MC engine:
// counter-based random numbers, the decision exp, move proposals, rigid
// placement and the receptor-pair energy kernel body.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/geauxdock.h"

namespace gd {

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11). Replaces the per-thread
// cuRAND XORWOW states of the reference (src/kernel_cuda_l1_initcurand.cu:4-16,
// src/kernel_cuda_l3_util.cu:115-126): a draw is a pure function of
// (seed, complex, replica, step, purpose), independent of launch geometry.
// ---------------------------------------------------------------------------
struct u32x4 {
  uint32_t x, y, z, w;
};

__host__ __device__ inline u32x4 philox4x32_10(u32x4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r > 0) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    const uint64_t p0 = (uint64_t)0xD2511F53u * c.x;
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c.z;
    const u32x4 n = {(uint32_t)(p1 >> 32) ^ c.y ^ k0, (uint32_t)p1,
                     (uint32_t)(p0 >> 32) ^ c.w ^ k1, (uint32_t)p0};
    c = n;
  }
  return c;
}

// (0, 1]: 24 random bits plus one, times 2^-24 (exact in float).
__host__ __device__ inline float uniform01(uint32_t b) {
  return (float)((b >> 8) + 1u) * 5.9604644775390625e-08f;
}

// Stream layout: MC step of replica `flat` of complex `cid` uses counters
// (step, flat, cid, 0) and (step, flat, cid, 1) -> eight uniforms; moves use
// u[0..5], the Metropolis test u[6], the delayed-acceptance second stage u[7].
// Exchange pair `pair` of ladder `group` at segment `seg`: (seg, group, cid,
// 0x100 + pair), first word.
__device__ inline void mc_uniforms(uint32_t k0, uint32_t k1, int cid, int flat, int step,
                                   float u[8]) {
  const u32x4 a = philox4x32_10({(uint32_t)step, (uint32_t)flat, (uint32_t)cid, 0u}, k0, k1);
  const u32x4 b = philox4x32_10({(uint32_t)step, (uint32_t)flat, (uint32_t)cid, 1u}, k0, k1);
  u[0] = uniform01(a.x);
  u[1] = uniform01(a.y);
  u[2] = uniform01(a.z);
  u[3] = uniform01(a.w);
  u[4] = uniform01(b.x);
  u[5] = uniform01(b.y);
  u[6] = uniform01(b.z);
  u[7] = uniform01(b.w);
}

__device__ inline float exchange_uniform(uint32_t k0, uint32_t k1, int cid, int group, int seg,
                                         int pair) {
  const u32x4 a = philox4x32_10(
      {(uint32_t)seg, (uint32_t)group, (uint32_t)cid, 0x100u + (uint32_t)pair}, k0, k1);
  return uniform01(a.x);
}

// ---------------------------------------------------------------------------
// exp for Metropolis and exchange decisions. x = n ln2 + r, n = rint(x log2 e),
// Cody-Waite two-constant reduction (LN2_HI carries 9 trailing zero bits), e^r
// by its degree-7 Taylor polynomial in Horner form with fused multiply-adds,
// then scaling by 2^n. Arguments above 88 give +inf, below -87 give 0, so
// every result is a normal number. Written with explicit fmaf so the value is
// the same bit pattern wherever it is evaluated.
// ---------------------------------------------------------------------------
__host__ __device__ inline float decision_expf(float x) {
  if (x != x) return x;
  if (x > 88.0f) return __builtin_inff();
  if (x < -87.0f) return 0.0f;
  const float n = rintf(x * 1.44269502162933349609375f);
  float r = fmaf(-n, 0.693145751953125f, x);
  r = fmaf(-n, 1.428606765330187045e-06f, r);
  float p = 1.98412698412698412698e-04f;
  p = fmaf(p, r, 1.38888888888888888889e-03f);
  p = fmaf(p, r, 8.33333333333333333333e-03f);
  p = fmaf(p, r, 4.16666666666666666667e-02f);
  p = fmaf(p, r, 1.66666666666666666667e-01f);
  p = fmaf(p, r, 0.5f);
  p = fmaf(p, r, 1.0f);
  p = fmaf(p, r, 1.0f);
  return ldexpf(p, (int)n);
}

// Accept_d (src/kernel_cuda_l2_accept.cu:23): U < exp(dE * minus_beta).
__device__ inline bool metropolis(float e_new, float e_old, float minus_beta, float u) {
  const float de = e_new - e_old;
  const float prod = de * minus_beta;
  return u < decision_expf(prod);
}

// Six DOF offsets of one step. Products and sums round one at a time, as in
// the CPU chain: no multiply-add fuses with the caller's addition, so the
// offsets and the moved pose have the same bits wherever they are inlined.
__device__ inline void move_offsets(const float u[8], const float scale[6], int kind,
                                    float dm[6]) {
#pragma clang fp contract(off)
  if (kind == GD_MOVE_GAUSSIAN) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const float r = sqrtf(-2.0f * logf(u[2 * k]));
      float s, c;
      sincosf(6.28318530717958647692f * u[2 * k + 1], &s, &c);
      dm[2 * k] = scale[2 * k] * (r * c);
      dm[2 * k + 1] = scale[2 * k + 1] * (r * s);
    }
  } else {
#pragma unroll
    for (int i = 0; i < 6; ++i) dm[i] = scale[i] * (2.0f * u[i] - 1.0f);
  }
}

// The moved pose mn = mo + dm, each sum rounded on its own (see above).
__device__ inline void apply_move(const float mo[6], const float dm[6], float mn[6]) {
#pragma clang fp contract(off)
#pragma unroll
  for (int i = 0; i < 6; ++i) mn[i] = mo[i] + dm[i];
}

// Euler rotation of src/kernel_cuda_l2_move.cu:43-51.
struct Rot {
  float r00, r01, r02, r10, r11, r12, r20, r21, r22;
};

__device__ inline Rot euler(float a, float b, float g) {
  float s1, c1, s2, c2, s3, c3;
  sincosf(a, &s1, &c1);
  sincosf(b, &s2, &c2);
  sincosf(g, &s3, &c3);
  Rot R;
  R.r00 = c1 * c2;
  R.r01 = c1 * s2 * s3 - c3 * s1;
  R.r02 = s1 * s3 + c1 * c3 * s2;
  R.r10 = c2 * s1;
  R.r11 = c1 * c3 + s1 * s2 * s3;
  R.r12 = c3 * s1 * s2 - c1 * s3;
  R.r20 = -s2;
  R.r21 = c2 * s3;
  R.r22 = c2 * c3;
  return R;
}

// Distance used by every contact decision (energy loop, contact map, MCC
// side loop): identical instruction sequence everywhere.
__device__ __forceinline__ float pair_d2(float lx, float ly, float lz, float px, float py,
                                         float pz) {
  const float dx = lx - px, dy = ly - py, dz = lz - pz;
  return fmaf(dx, dx, fmaf(dy, dy, dz * dz));
}

__device__ __forceinline__ float fast_rsq(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ float fast_rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float fast_exp2(float x) { return __builtin_amdgcn_exp2f(x); }

}  // namespace gd
