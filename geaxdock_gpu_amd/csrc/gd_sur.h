// gd_sur.h -- device building blocks of the learned energy surrogate: feature
// layout, radial features and their grids, BF16 MFMA MLP tiles (FP32
// accumulate). The kernels are in gd_sur.hip; the MC kernels use the tiles
// for the direct network evaluation.
//
// Per ligand atom: 56 features (24 one-hot ligand type, 32 radial counts of
// receptor points by class t mod 4 at 8 radial centres 1..8 A, Gaussian width
// 1 A, smooth cosine cutoff at 8 A), plus a constant 1 in column 56 that
// carries the first-layer bias. MLP 56 -> 128 -> 128 -> 1 with SiLU; a pose's
// value is the sum over its atoms. Rows are padded to 64 columns (bf16).
//
// MFMA: v_mfma_f32_16x16x32_bf16 (gfx950) on wave64, C[16x16] += A[16x32]
// B[32x16]. Fragment layout used throughout (lane l, i = 0..7, j = 0..3):
//   A[16x32]: A[l % 16][8 (l / 16) + i]      (row-major run of 8)
//   B[32x16]: B[8 (l / 16) + i][l % 16]      (column-major run of 8)
//   C[16x16]: C[4 (l / 16) + j][l % 16]
#pragma once

#include "gd_device.h"
#include "gd_kernels.h"

namespace gd {

constexpr int SUR_F = 64;        // padded feature width (56 + bias column + zeros)
constexpr int SUR_NF = 56;       // real features
constexpr int SUR_BIAS_COL = 56; // constant-1 column
constexpr int SUR_H = 128;
constexpr int SUR_NCLS = 4, SUR_NRBF = 8;
constexpr float SUR_CUT = 8.0f;
// parameter vector (FP32): W1[64][128] (row 56 = b1), W2[128][128], b2[128], w3[128], b3
constexpr int SP_W1 = 0;
constexpr int SP_W2 = SP_W1 + SUR_F * SUR_H;
constexpr int SP_B2 = SP_W2 + SUR_H * SUR_H;
constexpr int SP_W3 = SP_B2 + SUR_H;
constexpr int SP_B3 = SP_W3 + SUR_H;
constexpr int SUR_NPARAM = SP_B3 + 1;

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4_t;

__host__ __device__ inline unsigned short f2bf(float f) {
  unsigned int u;
  memcpy(&u, &f, 4);
  u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even
  return (unsigned short)(u >> 16);
}
__device__ inline float bf2f(unsigned short h) { return __uint_as_float((unsigned int)h << 16); }

__device__ inline float silu(float x) { return x * fast_rcp(1.0f + fast_exp2(-x * 1.44269504f)); }
__device__ inline float dsilu(float x) {
  const float s = fast_rcp(1.0f + fast_exp2(-x * 1.44269504f));
  return s * (1.0f + x * (1.0f - s));
}

__device__ inline f32x4_t mfma32(bf16x8_t a, bf16x8_t b, f32x4_t c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}
// eight bf16 at p (16-byte aligned)
__device__ inline bf16x8_t ld8(const unsigned short *p) {
  return *reinterpret_cast<const bf16x8_t *>(p);
}

// ---------------------------------------------------------------------------
// Radial features of one position against a receptor conformer (sorted
// points with per-point class).
// ---------------------------------------------------------------------------
__device__ inline void radial32(float x, float y, float z, const float4 *xyze, const int *cls,
                                int pnp, float (&f)[32]) {
#pragma unroll
  for (int i = 0; i < 32; ++i) f[i] = 0.0f;
  const float c2 = SUR_CUT * SUR_CUT;
  for (int p = 0; p < pnp; ++p) {
    const float4 P = xyze[p];
    const float d2 = pair_d2(x, y, z, P.x, P.y, P.z);
    if (d2 < c2) {
      const float d = sqrtf(d2);
      const float fc = 0.5f * (1.0f + cosf(3.14159265358979f * d / SUR_CUT));
      const int c = cls[p];
#pragma unroll
      for (int k = 0; k < SUR_NRBF; ++k) {
        const float t = d - (float)(k + 1);
        const float g = expf(-0.5f * t * t) * fc;
#pragma unroll
        for (int cc = 0; cc < SUR_NCLS; ++cc)
          if (cc == c) f[cc * SUR_NRBF + k] += g;
      }
    }
  }
}

struct SurGrid {
  const float *grid;  // [n_prt][n^3][32]
  float ox, oy, oz, h;
  int n;
};

// Trilinear interpolation of the 32 radial features at (x, y, z), clamped.
__device__ inline void grid_radial(const SurGrid &G, int prt, float x, float y, float z,
                                   float (&f)[32]) {
  const int n = G.n;
  const float lim = (float)(n - 1) - 1e-4f;
  const float ux = fminf(fmaxf((x - G.ox) / G.h, 0.0f), lim);
  const float uy = fminf(fmaxf((y - G.oy) / G.h, 0.0f), lim);
  const float uz = fminf(fmaxf((z - G.oz) / G.h, 0.0f), lim);
  const int ix = (int)ux, iy = (int)uy, iz = (int)uz;
  const float tx = ux - ix, ty = uy - iy, tz = uz - iz;
  const float *base = G.grid + (size_t)prt * n * n * n * 32;
#pragma unroll
  for (int q = 0; q < 32; ++q) f[q] = 0.0f;
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    const int dx = c >> 2, dy = (c >> 1) & 1, dz = c & 1;
    const float w = (dx ? tx : 1.0f - tx) * (dy ? ty : 1.0f - ty) * (dz ? tz : 1.0f - tz);
    const float4 *src = reinterpret_cast<const float4 *>(
        base + (((size_t)(ix + dx) * n + (iy + dy)) * n + (iz + dz)) * 32);
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const float4 v = src[q];
      f[4 * q] = fmaf(w, v.x, f[4 * q]);
      f[4 * q + 1] = fmaf(w, v.y, f[4 * q + 1]);
      f[4 * q + 2] = fmaf(w, v.z, f[4 * q + 2]);
      f[4 * q + 3] = fmaf(w, v.w, f[4 * q + 3]);
    }
  }
}

// Feature row of ligand atom l of one pose (movematrix m, ligand conformer
// il, protein conformer ip): bf16 into X [row][64] and, when xt != null, the
// transpose XT [64][ld]; FP32 features into xf when given.
__device__ inline void feat_row(const DevComplex &C, const SurGrid &G, const float *m, int il,
                                int ip, int l, int64_t row, unsigned short *X,
                                unsigned short *XT, int64_t ld, float *xf) {
  const int lna = C.lna;
  const Rot R = euler(m[3], m[4], m[5]);
  const float *orig = C.lig_orig + (size_t)il * 3 * lna;
  const float ox = orig[l], oy = orig[lna + l], oz = orig[2 * lna + l];
  const float x = R.r00 * ox + R.r01 * oy + R.r02 * oz + m[0] + C.pocket[0];
  const float y = R.r10 * ox + R.r11 * oy + R.r12 * oz + m[1] + C.pocket[1];
  const float z = R.r20 * ox + R.r21 * oy + R.r22 * oz + m[2] + C.pocket[2];
  float f[32];
  grid_radial(G, ip, x, y, z, f);
  const int lt = C.lig_t[l];
  float v[SUR_F];
#pragma unroll
  for (int q = 0; q < SUR_F; ++q) v[q] = 0.0f;
#pragma unroll
  for (int q = 0; q < NLT; ++q) v[q] = (q == lt) ? 1.0f : 0.0f;
#pragma unroll
  for (int q = 0; q < 32; ++q) v[NLT + q] = f[q];
  v[SUR_BIAS_COL] = 1.0f;
  if (X)
#pragma unroll
    for (int q = 0; q < SUR_F; ++q) X[row * SUR_F + q] = f2bf(v[q]);
  if (XT)
#pragma unroll
    for (int q = 0; q < SUR_F; ++q) XT[(int64_t)q * ld + row] = f2bf(v[q]);
  if (xf)
#pragma unroll
    for (int q = 0; q < SUR_NF; ++q) xf[q] = v[q];
}

// bf16 weight copies in the layouts the fragments read:
//   W1T[128][64] = W1^T, W2T[128][128] = W2^T, W2R[128][128] = W2.
struct SurW {
  const unsigned short *W1T, *W2T, *W2R;
  const float *b2, *w3, *b3;
};

// Forward of one wave's 16 rows. Returns h1, h2 pre-activations in C
// layout; writes bf16 A1 rows into the wave's LDS tile a1s[16][128].
__device__ inline void mlp_fwd16(const unsigned short *X, int64_t r0, const SurW &W,
                                 unsigned short *a1s, f32x4_t (&h1)[8], f32x4_t (&h2)[8]) {
  const int l = threadIdx.x & 63;
  const int lr = l & 15, lq = l >> 4;
  bf16x8_t xa[2];
#pragma unroll
  for (int ks = 0; ks < 2; ++ks) xa[ks] = ld8(X + (r0 + lr) * SUR_F + 32 * ks + 8 * lq);
#pragma unroll
  for (int t = 0; t < 8; ++t) {
    f32x4_t c = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int ks = 0; ks < 2; ++ks)
      c = mfma32(xa[ks], ld8(W.W1T + (16 * t + lr) * SUR_F + 32 * ks + 8 * lq), c);
    h1[t] = c;
#pragma unroll
    for (int i = 0; i < 4; ++i) a1s[(4 * lq + i) * SUR_H + 16 * t + lr] = f2bf(silu(c[i]));
  }
  __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0): LDS writes visible to the wave
  __builtin_amdgcn_wave_barrier();
#pragma unroll
  for (int u = 0; u < 8; ++u) {
    f32x4_t c = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int ks = 0; ks < 4; ++ks)
      c = mfma32(ld8(a1s + lr * SUR_H + 32 * ks + 8 * lq),
                 ld8(W.W2T + (16 * u + lr) * SUR_H + 32 * ks + 8 * lq), c);
    const float bb = W.b2[16 * u + lr];
#pragma unroll
    for (int i = 0; i < 4; ++i) c[i] += bb;
    h2[u] = c;
  }
}

// y for the wave's 16 rows: lane with l%16 == 0 holds rows 4*(l/16)+i.
__device__ inline void mlp_out16(const f32x4_t (&h2)[8], const SurW &W, float (&y)[4]) {
  const int l = threadIdx.x & 63;
  const int lr = l & 15;
#pragma unroll
  for (int i = 0; i < 4; ++i) y[i] = 0.0f;
#pragma unroll
  for (int u = 0; u < 8; ++u) {
    const float w = W.w3[16 * u + lr];
#pragma unroll
    for (int i = 0; i < 4; ++i) y[i] = fmaf(silu(h2[u][i]), w, y[i]);
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    float v = y[i];
    v += __shfl_xor(v, 1);
    v += __shfl_xor(v, 2);
    v += __shfl_xor(v, 4);
    v += __shfl_xor(v, 8);
    y[i] = v + W.b3[0];
  }
}

}  // namespace gd
