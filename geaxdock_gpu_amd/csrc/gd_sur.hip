// gd_sur.hip -- learned energy surrogate kernels: receptor feature grids,
// feature rows, BF16 MFMA MLP forward/backward (FP32 accumulate), fixed-order
// gradient sums, FP32 Adam. Building blocks and layouts: gd_sur.h.
#include "gd_sur.h"

namespace gd {

// Feature grid of one protein conformer: n^3 nodes at origin + h*(i,j,k),
// 32 floats per node, node index (i*n + j)*n + k.
extern "C" __global__ void gd_sur_grid_kernel(const float4 *xyze, const int *cls, int pnp,
                                              float ox, float oy, float oz, float h, int n,
                                              float *grid) {
  const int64_t node = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t total = (int64_t)n * n * n;
  if (node >= total) return;
  const int k = (int)(node % n);
  const int j = (int)((node / n) % n);
  const int i = (int)(node / ((int64_t)n * n));
  float f[32];
  radial32(ox + h * i, oy + h * j, oz + h * k, xyze, cls, pnp, f);
  float4 *o = reinterpret_cast<float4 *>(grid + node * 32);
#pragma unroll
  for (int q = 0; q < 8; ++q) o[q] = make_float4(f[4 * q], f[4 * q + 1], f[4 * q + 2], f[4 * q + 3]);
}

// Feature rows of poses of one complex: one thread per ligand atom of each
// pose, rows from row0 on.
extern "C" __global__ void gd_sur_feat_kernel(const DevComplex *cxs, int cxi, SurGrid G, int n,
                                              const int *idx_lig, const int *idx_prt,
                                              const float *mm, int64_t row0,
                                              unsigned short *X, unsigned short *XT,
                                              int64_t ld, float *xf) {
  const DevComplex &C = cxs[cxi];
  const int lna = C.lna;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)n * lna) return;
  const int pose = (int)(t / lna), l = (int)(t % lna);
  feat_row(C, G, mm + (size_t)pose * 6, idx_lig[pose], idx_prt[pose], l, row0 + t, X, XT, ld,
           xf ? xf + (size_t)t * SUR_NF : nullptr);
}

// Feature rows of poses of many complexes: one workgroup per pose, its rows
// from row_of[pose] on; the grid geometry comes from the complex record.
extern "C" __global__ void gd_sur_feat_multi_kernel(const DevComplex *cxs, int n, const int *cid,
                                                    const int *idx_lig, const int *idx_prt,
                                                    const float *mm, const int64_t *row_of,
                                                    unsigned short *X, unsigned short *XT,
                                                    int64_t ld) {
  const int pose = blockIdx.x;
  if (pose >= n) return;
  const DevComplex &C = cxs[cid[pose]];
  SurGrid G;
  G.grid = C.sgrid;
  G.ox = C.sox;
  G.oy = C.soy;
  G.oz = C.soz;
  G.h = C.sh;
  G.n = C.sn;
  for (int l = threadIdx.x; l < C.lna; l += blockDim.x)
    feat_row(C, G, mm + (size_t)pose * 6, idx_lig[pose], idx_prt[pose], l, row_of[pose] + l, X, XT,
             ld, nullptr);
}

// Per-row outputs of a batch: 4 waves x 16 rows per workgroup.
extern "C" __global__ void __launch_bounds__(256) gd_mlp_fwd_kernel(const unsigned short *X,
                                                                    int64_t n_rows, SurW W,
                                                                    float *y) {
  __shared__ __attribute__((aligned(16))) unsigned short a1s[4][16 * SUR_H];
  const int wv = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int64_t r0 = ((int64_t)blockIdx.x * 4 + wv) * 16;
  if (r0 >= n_rows) return;  // whole wave leaves; no block barrier below
  f32x4_t h1[8], h2[8];
  mlp_fwd16(X, r0, W, a1s[wv], h1, h2);
  float yy[4];
  mlp_out16(h2, W, yy);
  if ((l & 15) == 0)
#pragma unroll
    for (int i = 0; i < 4; ++i) y[r0 + 4 * (l >> 4) + i] = yy[i];
}

// Network output at every grid point for each listed ligand atom type:
// Y[type][point] = MLP(onehot(type), features(point)), the features being
// the grid's own values (no interpolation). One wave per 16 points of one
// type; 4 waves per workgroup.
extern "C" __global__ void __launch_bounds__(256) gd_sur_tab_kernel(const float *fgrid,
                                                                    int n_pts,
                                                                    const int *types,
                                                                    int n_types, SurW W,
                                                                    float *Y) {
  __shared__ __attribute__((aligned(16))) unsigned short xs[4][16 * SUR_F];
  __shared__ __attribute__((aligned(16))) unsigned short a1s[4][16 * SUR_H];
  const int wv = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int tiles = (n_pts + 15) / 16;
  const int64_t job = (int64_t)blockIdx.x * 4 + wv;
  if (job >= (int64_t)tiles * n_types) return;  // whole wave leaves; no block barrier
  const int ti = (int)(job / tiles);
  const int p0 = (int)(job % tiles) * 16;
  const int type = types[ti];
  // lane l fills columns 16 (l % 4) .. +15 of row l / 4
  {
    const int r = l >> 2, c0 = 16 * (l & 3);
    const int pt = min(p0 + r, n_pts - 1);
    const float *f = fgrid + (size_t)pt * 32;
    unsigned short *row = xs[wv] + r * SUR_F;
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int c = c0 + q;
      float v = 0.0f;
      if (c < NLT) v = (c == type) ? 1.0f : 0.0f;
      else if (c < NLT + 32) v = f[c - NLT];
      else if (c == SUR_BIAS_COL) v = 1.0f;
      row[c] = f2bf(v);
    }
  }
  __builtin_amdgcn_s_waitcnt(0xc07f);
  __builtin_amdgcn_wave_barrier();
  f32x4_t h1[8], h2[8];
  mlp_fwd16(xs[wv], 0, W, a1s[wv], h1, h2);
  float yy[4];
  mlp_out16(h2, W, yy);
  if ((l & 15) == 0)
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int pt = p0 + 4 * (l >> 4) + i;
      if (pt < n_pts) Y[(size_t)type * n_pts + pt] = yy[i];
    }
}

// Pose sums over poses [p0, p1): S = sum of row outputs, residual against the
// target, dS = 2 (S - target) / (p1 - p0), squared residuals for the loss.
// Row offsets are absolute; y is indexed by absolute row.
extern "C" __global__ void gd_mlp_pose_kernel(const float *y, const int64_t *off, int p0, int p1,
                                              const float *target, float *dS, float *sq,
                                              float *S_out) {
  const int p = p0 + blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= p1) return;
  float s = 0.0f;
  for (int64_t r = off[p]; r < off[p + 1]; ++r) s += y[r];
  const float res = s - target[p];
  dS[p] = 2.0f * res / (float)(p1 - p0);
  sq[p] = res * res;
  if (S_out) S_out[p] = s;
}

// Row stride (bf16) of the transposed LDS tiles of gd_mlp_train_kernel: 64
// rows and 8 of padding, 144 bytes (a multiple of 16 for 16-byte reads).
constexpr int SUR_TR = 72;
constexpr int SUR_STEP = 64;  // rows per step of gd_mlp_train_kernel (4 waves x 16)

// Backward of one batch fused with its weight gradients. Workgroup b owns
// rows [b R, (b + 1) R) (R a multiple of 64) and walks them 64 at a time;
// wave w recomputes the forward of rows 16 w .. 16 w + 15 of the step, forms
//   dy = dS of the row's pose,  dH2 = dy w3 * silu'(H2),
//   dA1 = dH2 W2^T,             dH1 = dA1 * silu'(H1)
// and writes X, A1 = silu(H1), dH2 and dH1 of its rows transposed into LDS.
// After a workgroup barrier wave w accumulates its share of the step's
//   dW2 += A1^T dH2   (row tiles 2w and 2w + 1 of 8),
//   dW1 += X^T dH1    (row tile w of 4; row 56 is the b1 gradient)
// in MFMA accumulators. Nothing but the partials goes back to HBM, in a fixed
// order (gd_mlp_gradsum_kernel and gd_mlp_w3sum_kernel sum them, so training
// is deterministic): p1g[b][64][128], p2g[b][128][128], pb2[b][128] and
// pw3[4 b + w][129] (w3, then b3).
extern "C" __global__ void __launch_bounds__(256) gd_mlp_train_kernel(
    const unsigned short *X, int64_t n_rows, int R, const int *pose_of_row, int p0, int p1,
    const float *dS, SurW W, float *p1g, float *p2g, float *pb2, float *pw3) {
  extern __shared__ __attribute__((aligned(16))) unsigned char sm_train[];
  unsigned short *a1t = reinterpret_cast<unsigned short *>(sm_train);  // A1^T [128][SUR_TR]
  unsigned short *h2t = a1t + SUR_H * SUR_TR;                          // dH2^T
  unsigned short *h1t = h2t + SUR_H * SUR_TR;                          // dH1^T
  unsigned short *xt = h1t + SUR_H * SUR_TR;                           // X^T [64][SUR_TR]
  unsigned short *rows = xt + SUR_F * SUR_TR;                          // [4][16][128] row-major
  float *red = reinterpret_cast<float *>(rows + 4 * 16 * SUR_H);      // [4][128]
  const int wv = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int lr = l & 15, lq = l >> 4;
  const int64_t base = (int64_t)blockIdx.x * R;
  const int64_t end = base + R < n_rows ? base + R : n_rows;
  unsigned short *ts = rows + wv * 16 * SUR_H;
  f32x4_t g2[2][8], g1[8];
#pragma unroll
  for (int n = 0; n < 8; ++n) {
    g2[0][n] = g2[1][n] = g1[n] = f32x4_t{0.f, 0.f, 0.f, 0.f};
  }
  float gw3[8], gb2[8], gb3 = 0.0f;
#pragma unroll
  for (int u = 0; u < 8; ++u) gw3[u] = gb2[u] = 0.0f;

  for (int64_t s0 = base; s0 < end; s0 += SUR_STEP) {
    const int64_t r0 = s0 + 16 * wv;
    const int cw = 16 * wv;  // this wave's columns in the transposed tiles
    f32x4_t h1[8], h2[8];
    mlp_fwd16(X, r0, W, ts, h1, h2);
    // X^T: lane (lr, lq) moves columns 16 lq .. 16 lq + 15 of row lr
    {
      const unsigned short *xr = X + (r0 + lr) * SUR_F + 16 * lq;
#pragma unroll
      for (int q = 0; q < 16; ++q) xt[(16 * lq + q) * SUR_TR + cw + lr] = xr[q];
    }
    // A1^T from the C layout (rows 4 lq + i, column 16 t + lr)
#pragma unroll
    for (int t = 0; t < 8; ++t) {
      unsigned short v[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) v[i] = f2bf(silu(h1[t][i]));
      *reinterpret_cast<uint2 *>(a1t + (16 * t + lr) * SUR_TR + cw + 4 * lq) =
          make_uint2(v[0] | ((unsigned)v[1] << 16), v[2] | ((unsigned)v[3] << 16));
    }
    float dy[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int p = pose_of_row[r0 + 4 * lq + i];
      dy[i] = (p >= p0 && p < p1) ? dS[p] : 0.0f;
    }
    // w3 and b3 (rows counted once: b3 from the lr == 0 lanes)
#pragma unroll
    for (int u = 0; u < 8; ++u)
#pragma unroll
      for (int i = 0; i < 4; ++i) gw3[u] = fmaf(silu(h2[u][i]), dy[i], gw3[u]);
    if (lr == 0) gb3 += (dy[0] + dy[1]) + (dy[2] + dy[3]);
    // dH2: row-major into the wave's tile (A operand of dA1) and transposed
    __builtin_amdgcn_s_waitcnt(0xc07f);  // the forward's LDS reads of ts are done
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const float w = W.w3[16 * u + lr];
      unsigned short v[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const float g = dy[i] * w * dsilu(h2[u][i]);
        gb2[u] += g;
        v[i] = f2bf(g);
        ts[(4 * lq + i) * SUR_H + 16 * u + lr] = v[i];
      }
      *reinterpret_cast<uint2 *>(h2t + (16 * u + lr) * SUR_TR + cw + 4 * lq) =
          make_uint2(v[0] | ((unsigned)v[1] << 16), v[2] | ((unsigned)v[3] << 16));
    }
    __builtin_amdgcn_s_waitcnt(0xc07f);
    __builtin_amdgcn_wave_barrier();
    // dA1 = dH2 W2^T (B[k][n] = W2[n][k] = W2R[n][k]), dH1 transposed
#pragma unroll
    for (int t = 0; t < 8; ++t) {
      f32x4_t c = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int ks = 0; ks < 4; ++ks)
        c = mfma32(ld8(ts + lr * SUR_H + 32 * ks + 8 * lq),
                   ld8(W.W2R + (16 * t + lr) * SUR_H + 32 * ks + 8 * lq), c);
      unsigned short v[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) v[i] = f2bf(c[i] * dsilu(h1[t][i]));
      *reinterpret_cast<uint2 *>(h1t + (16 * t + lr) * SUR_TR + cw + 4 * lq) =
          make_uint2(v[0] | ((unsigned)v[1] << 16), v[2] | ((unsigned)v[3] << 16));
    }
    __syncthreads();
    // weight gradients over the step's 64 rows (two K blocks of 32)
#pragma unroll
    for (int kb = 0; kb < 2; ++kb) {
      const int ko = 32 * kb + 8 * lq;
#pragma unroll
      for (int a = 0; a < 2; ++a) {
        const bf16x8_t A = ld8(a1t + (16 * (2 * wv + a) + lr) * SUR_TR + ko);
#pragma unroll
        for (int n = 0; n < 8; ++n)
          g2[a][n] = mfma32(A, ld8(h2t + (16 * n + lr) * SUR_TR + ko), g2[a][n]);
      }
      const bf16x8_t A = ld8(xt + (16 * wv + lr) * SUR_TR + ko);
#pragma unroll
      for (int n = 0; n < 8; ++n) g1[n] = mfma32(A, ld8(h1t + (16 * n + lr) * SUR_TR + ko), g1[n]);
    }
    __syncthreads();  // the tiles are rewritten by the next step
  }

  // partials of this workgroup
  float *o2 = p2g + (int64_t)blockIdx.x * SUR_H * SUR_H;
  float *o1 = p1g + (int64_t)blockIdx.x * SUR_F * SUR_H;
#pragma unroll
  for (int n = 0; n < 8; ++n)
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      o2[(16 * (2 * wv) + 4 * lq + i) * SUR_H + 16 * n + lr] = g2[0][n][i];
      o2[(16 * (2 * wv + 1) + 4 * lq + i) * SUR_H + 16 * n + lr] = g2[1][n][i];
      o1[(16 * wv + 4 * lq + i) * SUR_H + 16 * n + lr] = g1[n][i];
    }
  // w3, b3: sum over the lane quarters (rows), one partial row per wave
#pragma unroll
  for (int u = 0; u < 8; ++u) {
    float v = gw3[u];
    v += __shfl_xor(v, 16);
    v += __shfl_xor(v, 32);
    gw3[u] = v;
    float b = gb2[u];
    b += __shfl_xor(b, 16);
    b += __shfl_xor(b, 32);
    gb2[u] = b;
  }
  gb3 += __shfl_xor(gb3, 16);
  gb3 += __shfl_xor(gb3, 32);
  float *w3o = pw3 + ((int64_t)blockIdx.x * 4 + wv) * (SUR_H + 1);
  if (lq == 0) {
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      w3o[16 * u + lr] = gw3[u];
      red[wv * SUR_H + 16 * u + lr] = gb2[u];
    }
    if (lr == 0) w3o[SUR_H] = gb3;
  }
  __syncthreads();
  // b2: the four waves' column sums in a fixed order
  if (threadIdx.x < SUR_H) {
    const int j = threadIdx.x;
    pb2[(int64_t)blockIdx.x * SUR_H + j] =
        (red[j] + red[SUR_H + j]) + (red[2 * SUR_H + j] + red[3 * SUR_H + j]);
  }
}

// Dynamic LDS bytes of gd_mlp_train_kernel.
constexpr int SUR_TRAIN_LDS = (3 * SUR_H * SUR_TR + SUR_F * SUR_TR + 4 * 16 * SUR_H) * 2 +
                              4 * SUR_H * 4;

// Sum of the squared residuals of poses [0, n) into *out: one workgroup,
// strided partial sums then a fixed-shape tree (deterministic).
extern "C" __global__ void __launch_bounds__(256) gd_mlp_sqsum_kernel(const float *sq, int n,
                                                                     float *out) {
  __shared__ float red[256];
  float s = 0.0f;
  for (int i = threadIdx.x; i < n; i += 256) s += sq[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = red[0];
}

// Sum of n partials part[c * stride + q], c = 0..n-1, in a fixed order:
// eight interleaved accumulators (eight loads in flight) then a fixed tree.
__device__ inline float sum_partials(const float *part, int n, int64_t stride, int q) {
  float a[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  int c = 0;
  for (; c + 8 <= n; c += 8)
#pragma unroll
    for (int j = 0; j < 8; ++j) a[j] += part[(int64_t)(c + j) * stride + q];
  for (int j = 0; c < n; ++c, ++j) a[j] += part[(int64_t)c * stride + q];
  return ((a[0] + a[1]) + (a[2] + a[3])) + ((a[4] + a[5]) + (a[6] + a[7]));
}

// Sums gradient partials into g[SUR_NPARAM] (fixed order: deterministic).
extern "C" __global__ void gd_mlp_gradsum_kernel(const float *p1, int n1, const float *p2, int n2,
                                                 const float *pb2, const float *pw3, int64_t nw,
                                                 float *g) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= SUR_NPARAM) return;
  float s = 0.0f;
  if (q < SP_W2) {
    s = sum_partials(p1, n1, (int64_t)SUR_F * SUR_H, q);
  } else if (q < SP_B2) {
    s = sum_partials(p2, n2, (int64_t)SUR_H * SUR_H, q - SP_W2);
  } else if (q < SP_W3) {
    s = sum_partials(pb2, n2, SUR_H, q - SP_B2);
  } else {
    return;  // w3 and b3: gd_mlp_w3sum_kernel
  }
  g[q] = s;
}

// w3[j] and b3 (j == 128) gradients from the per-wave partials: one
// workgroup per output, strided partial sums then a fixed-shape tree
// (deterministic).
extern "C" __global__ void __launch_bounds__(256) gd_mlp_w3sum_kernel(const float *pw3, int64_t nw,
                                                                      float *g) {
  __shared__ float red[256];
  const int j = blockIdx.x;
  float s = 0.0f;
  for (int64_t w = threadIdx.x; w < nw; w += 256) s += pw3[w * (SUR_H + 1) + j];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) g[SP_W3 + j] = red[0];
}

// Adam on the FP32 master parameters; refreshes the bf16 layouts.
extern "C" __global__ void gd_mlp_adam_kernel(float *p, float *m, float *v, const float *g,
                                              float lr, float b1, float b2, float eps, float c1,
                                              float c2, unsigned short *W1T, unsigned short *W2T,
                                              unsigned short *W2R) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= SUR_NPARAM) return;
  float x = p[q];
  if (lr > 0.0f) {
    const float gq = g[q];
    const float mq = b1 * m[q] + (1.0f - b1) * gq;
    const float vq = b2 * v[q] + (1.0f - b2) * gq * gq;
    m[q] = mq;
    v[q] = vq;
    x -= lr * (mq / c1) / (sqrtf(vq / c2) + eps);
    p[q] = x;
  }
  if (q < SP_W2) {
    const int f = q / SUR_H, i = q % SUR_H;
    W1T[i * SUR_F + f] = f2bf(x);
  } else if (q < SP_B2) {
    const int i = (q - SP_W2) / SUR_H, j = (q - SP_W2) % SUR_H;
    W2T[j * SUR_H + i] = f2bf(x);
    W2R[i * SUR_H + j] = f2bf(x);
  }
}

}  // namespace gd
