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
  __shared__ unsigned short a1s[4][16 * SUR_H];
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
  __shared__ unsigned short xs[4][16 * SUR_F];
  __shared__ unsigned short a1s[4][16 * SUR_H];
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

// Backward of one batch: recomputes the forward, then per wave
//   dH2 = dy w3^T * silu'(H2);  dA1 = dH2 W2^T;  dH1 = dA1 * silu'(H1)
// and writes A1T[128][ld], dH2T[128][ld], dH1T[128][ld] (bf16, column per
// row) for the weight-gradient GEMMs, plus per-wave partials of grad w3, b3.
extern "C" __global__ void __launch_bounds__(256) gd_mlp_bwd_kernel(
    const unsigned short *X, int64_t n_rows, const int *pose_of_row, int p0, int p1,
    const float *dS, SurW W, unsigned short *A1T, unsigned short *dH2T, unsigned short *dH1T,
    int64_t ld, float *gw3_part) {
  __shared__ unsigned short tile[4][16 * SUR_H];
  const int wv = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int lr = l & 15, lq = l >> 4;
  const int64_t wave = (int64_t)blockIdx.x * 4 + wv;
  const int64_t r0 = wave * 16;
  if (r0 >= n_rows) return;
  unsigned short *ts = tile[wv];
  f32x4_t h1[8], h2[8];
  mlp_fwd16(X, r0, W, ts, h1, h2);
  // A1^T to global (rows r0 + 4 lq + i of column 16 t + lr)
#pragma unroll
  for (int t = 0; t < 8; ++t) {
    unsigned short v[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) v[i] = f2bf(silu(h1[t][i]));
    *reinterpret_cast<uint2 *>(A1T + (int64_t)(16 * t + lr) * ld + r0 + 4 * lq) =
        make_uint2(v[0] | ((unsigned)v[1] << 16), v[2] | ((unsigned)v[3] << 16));
  }
  float dy[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int p = pose_of_row[r0 + 4 * lq + i];
    dy[i] = (p >= p0 && p < p1) ? dS[p] : 0.0f;
  }
  // grad w3 (column 16 u + lr) and b3
  float gw[8];
#pragma unroll
  for (int u = 0; u < 8; ++u) {
    float s = 0.0f;
#pragma unroll
    for (int i = 0; i < 4; ++i) s = fmaf(silu(h2[u][i]), dy[i], s);
    s += __shfl_xor(s, 16);
    s += __shfl_xor(s, 32);
    gw[u] = s;
  }
  float gb = dy[0] + dy[1] + dy[2] + dy[3];
  gb += __shfl_xor(gb, 16);
  gb += __shfl_xor(gb, 32);
  if (lq == 0) {
#pragma unroll
    for (int u = 0; u < 8; ++u) gw3_part[wave * (SUR_H + 1) + 16 * u + lr] = gw[u];
    if (lr == 0) gw3_part[wave * (SUR_H + 1) + SUR_H] = gb;
  }
  // dH2 into the wave tile (row-major, for dA1) and to global (transposed)
  __builtin_amdgcn_wave_barrier();
#pragma unroll
  for (int u = 0; u < 8; ++u) {
    const float w = W.w3[16 * u + lr];
    unsigned short v[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      v[i] = f2bf(dy[i] * w * dsilu(h2[u][i]));
      ts[(4 * lq + i) * SUR_H + 16 * u + lr] = v[i];
    }
    *reinterpret_cast<uint2 *>(dH2T + (int64_t)(16 * u + lr) * ld + r0 + 4 * lq) =
        make_uint2(v[0] | ((unsigned)v[1] << 16), v[2] | ((unsigned)v[3] << 16));
  }
  __builtin_amdgcn_s_waitcnt(0xc07f);
  __builtin_amdgcn_wave_barrier();
  // dA1 = dH2 W2^T: B[j][i] = W2[i][j]
#pragma unroll
  for (int t = 0; t < 8; ++t) {
    f32x4_t c = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int ks = 0; ks < 8; ++ks)
      c = mfma16(ld4(ts + lr * SUR_H + 16 * ks + 4 * lq),
                 ld4(W.W2R + (16 * t + lr) * SUR_H + 16 * ks + 4 * lq), c);
    unsigned short v[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) v[i] = f2bf(c[i] * dsilu(h1[t][i]));
    *reinterpret_cast<uint2 *>(dH1T + (int64_t)(16 * t + lr) * ld + r0 + 4 * lq) =
        make_uint2(v[0] | ((unsigned)v[1] << 16), v[2] | ((unsigned)v[3] << 16));
  }
}

// Weight gradient partials G[M][128] = P^T Q over rows [c*R, (c+1)*R):
// P^T given as PT[M][ld] (M = 64 or 128), Q^T as QT[128][ld]. One workgroup
// (4 waves) per chunk; wave w owns output rows [16*mt, ...) for
// mt = w, w + 4, .... Column sums of Q (bias gradient) when qsum != null.
extern "C" __global__ void __launch_bounds__(256) gd_mlp_gemm_kernel(
    const unsigned short *PT, int64_t ldp, int M, const unsigned short *QT, int64_t ldq,
    int64_t n_rows, int R, float *part, float *qsum) {
  const int wv = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int lr = l & 15, lq = l >> 4;
  const int64_t b = (int64_t)blockIdx.x * R;
  const int64_t e = b + R < n_rows ? b + R : n_rows;
  const int mtiles = M / 16;
  float *out = part + (int64_t)blockIdx.x * M * SUR_H;
  for (int mt = wv; mt < mtiles; mt += 4) {
    f32x4_t c[8];
#pragma unroll
    for (int n = 0; n < 8; ++n) c[n] = {0.f, 0.f, 0.f, 0.f};
    float cs[8];
#pragma unroll
    for (int n = 0; n < 8; ++n) cs[n] = 0.0f;
    for (int64_t k0 = b; k0 < e; k0 += 16) {
      const bf16x4_t a = ld4(PT + (int64_t)(16 * mt + lr) * ldp + k0 + 4 * lq);
#pragma unroll
      for (int n = 0; n < 8; ++n) {
        const bf16x4_t q = ld4(QT + (int64_t)(16 * n + lr) * ldq + k0 + 4 * lq);
        c[n] = mfma16(a, q, c[n]);
        if (qsum && mt == 0)
#pragma unroll
          for (int i = 0; i < 4; ++i) cs[n] += bf2f((unsigned short)q[i]);
      }
    }
#pragma unroll
    for (int n = 0; n < 8; ++n)
#pragma unroll
      for (int i = 0; i < 4; ++i) out[(16 * mt + 4 * lq + i) * SUR_H + 16 * n + lr] = c[n][i];
    if (qsum && mt == 0) {
#pragma unroll
      for (int n = 0; n < 8; ++n) {
        float s = cs[n];
        s += __shfl_xor(s, 16);
        s += __shfl_xor(s, 32);
        if (lq == 0) qsum[(int64_t)blockIdx.x * SUR_H + 16 * n + lr] = s;
      }
    }
  }
}

// Sums gradient partials into g[SUR_NPARAM] (fixed order: deterministic).
extern "C" __global__ void gd_mlp_gradsum_kernel(const float *p1, int n1, const float *p2, int n2,
                                                 const float *pb2, const float *pw3, int64_t nw,
                                                 float *g) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= SUR_NPARAM) return;
  float s = 0.0f;
  if (q < SP_W2) {
    for (int c = 0; c < n1; ++c) s += p1[(int64_t)c * SUR_F * SUR_H + q];
  } else if (q < SP_B2) {
    for (int c = 0; c < n2; ++c) s += p2[(int64_t)c * SUR_H * SUR_H + (q - SP_W2)];
  } else if (q < SP_W3) {
    for (int c = 0; c < n2; ++c) s += pb2[(int64_t)c * SUR_H + (q - SP_B2)];
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
