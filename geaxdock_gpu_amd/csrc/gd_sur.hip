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

// Row strides (bf16) of the LDS tiles of the batch kernels: 64 + 8 and
// 128 + 8 (multiples of 8, so 16-byte reads stay aligned; the padding
// spreads a fragment's rows over the LDS banks).
constexpr int SUR_XS = SUR_F + 8;
constexpr int SUR_AS = SUR_H + 8;
constexpr int SUR_TR = 72;    // transposed tiles: 64 rows + 8
constexpr int SUR_STEP = 64;  // rows per step: 4 row tiles of 16

// Column-split MLP of the batch kernels. A workgroup of 4 waves takes 64
// rows at a time; wave w owns columns [32 w, 32 w + 32) of both 128-wide
// layers for all 64 rows, so its slices of W1, W2 and W2^T stay in
// registers for the whole kernel (B fragments: col tile c, K block kb), and
// the rows' X and A1 tiles are shared through LDS.
struct ColW {
  bf16x8_t w1[2][2];   // W1[:, 32w + 16c ..]: K = 64
  bf16x8_t w2[2][4];   // W2[:, 32w + 16c ..]: K = 128
  float b2[2], w3[2];  // per lane: column 32w + 16c + l % 16
};

__device__ inline void load_colw(const SurW &W, int wv, bool with_w2r, ColW &cw,
                                 bf16x8_t (&w2r)[2][4]) {
  const int l = threadIdx.x & 63, lr = l & 15, lq = l >> 4;
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    const int n = 32 * wv + 16 * c + lr;
#pragma unroll
    for (int kb = 0; kb < 2; ++kb) cw.w1[c][kb] = ld8(W.W1T + n * SUR_F + 32 * kb + 8 * lq);
#pragma unroll
    for (int kb = 0; kb < 4; ++kb) {
      cw.w2[c][kb] = ld8(W.W2T + n * SUR_H + 32 * kb + 8 * lq);
      if (with_w2r) w2r[c][kb] = ld8(W.W2R + n * SUR_H + 32 * kb + 8 * lq);
    }
    cw.b2[c] = W.b2[n];
    cw.w3[c] = W.w3[n];
  }
}

// X rows [r0, r0 + 64) as 512 pieces of 8 bf16, two per thread (all 256
// threads): read into registers, then written to xs (row-major) and, when xt
// is given, xt (transposed, [64 features][SUR_TR]).
__device__ inline void load_x(const unsigned short *X, int64_t r0, uint4 (&v)[2]) {
  const int t = threadIdx.x;
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int e = t + 256 * h;
    const int r = e >> 3, c8 = (e & 7) * 8;
    v[h] = *reinterpret_cast<const uint4 *>(X + (r0 + r) * SUR_F + c8);
  }
}

__device__ inline void put_x(const uint4 (&v)[2], unsigned short *xs, unsigned short *xt) {
  const int t = threadIdx.x;
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int e = t + 256 * h;
    const int r = e >> 3, c8 = (e & 7) * 8;
    *reinterpret_cast<uint4 *>(xs + r * SUR_XS + c8) = v[h];
    if (xt) {
      const unsigned int w[4] = {v[h].x, v[h].y, v[h].z, v[h].w};
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        xt[(c8 + 2 * q) * SUR_TR + r] = (unsigned short)(w[q] & 0xFFFFu);
        xt[(c8 + 2 * q + 1) * SUR_TR + r] = (unsigned short)(w[q] >> 16);
      }
    }
  }
}

__device__ inline void stage_x(const unsigned short *X, int64_t r0, unsigned short *xs,
                               unsigned short *xt) {
  uint4 v[2];
  load_x(X, r0, v);
  put_x(v, xs, xt);
}

// Layer 1 of the wave's column slice: h1[rt][c] = C tile (rows 16 rt.., cols
// 32w + 16c ..), from the staged X tile.
__device__ inline void col_layer1(const unsigned short *xs, const ColW &cw, f32x4_t (&h1)[4][2]) {
  const int l = threadIdx.x & 63, lr = l & 15, lq = l >> 4;
#pragma unroll
  for (int rt = 0; rt < 4; ++rt) h1[rt][0] = h1[rt][1] = f32x4_t{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int kb = 0; kb < 2; ++kb)
#pragma unroll
    for (int rt = 0; rt < 4; ++rt) {
      const bf16x8_t a = ld8(xs + (16 * rt + lr) * SUR_XS + 32 * kb + 8 * lq);
      h1[rt][0] = mfma32(a, cw.w1[0][kb], h1[rt][0]);
      h1[rt][1] = mfma32(a, cw.w1[1][kb], h1[rt][1]);
    }
}

// Layer 2 of the wave's column slice from the shared A1 tile (+ b2).
__device__ inline void col_layer2(const unsigned short *a1s, const ColW &cw, f32x4_t (&h2)[4][2]) {
  const int l = threadIdx.x & 63, lr = l & 15, lq = l >> 4;
#pragma unroll
  for (int rt = 0; rt < 4; ++rt) h2[rt][0] = h2[rt][1] = f32x4_t{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int kb = 0; kb < 4; ++kb)
#pragma unroll
    for (int rt = 0; rt < 4; ++rt) {
      const bf16x8_t a = ld8(a1s + (16 * rt + lr) * SUR_AS + 32 * kb + 8 * lq);
      h2[rt][0] = mfma32(a, cw.w2[0][kb], h2[rt][0]);
      h2[rt][1] = mfma32(a, cw.w2[1][kb], h2[rt][1]);
    }
#pragma unroll
  for (int rt = 0; rt < 4; ++rt)
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
      for (int i = 0; i < 4; ++i) h2[rt][c][i] += cw.b2[c];
}

// Store the C tiles v (rows 16 rt + 4 lq + i, column 32w + 16c + lr) as bf16:
// row-major into rm (row stride SUR_AS) and, when tr is given, transposed
// into tr ([128][SUR_TR]).
__device__ inline void put_tiles(const float (&v)[4][2][4], int wv, unsigned short *rm,
                                 unsigned short *tr) {
  const int l = threadIdx.x & 63, lr = l & 15, lq = l >> 4;
#pragma unroll
  for (int rt = 0; rt < 4; ++rt)
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      const int col = 32 * wv + 16 * c + lr;
      unsigned short b[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        b[i] = f2bf(v[rt][c][i]);
        if (rm) rm[(16 * rt + 4 * lq + i) * SUR_AS + col] = b[i];
      }
      if (tr)
        *reinterpret_cast<uint2 *>(tr + col * SUR_TR + 16 * rt + 4 * lq) =
            make_uint2(b[0] | ((unsigned)b[1] << 16), b[2] | ((unsigned)b[3] << 16));
    }
}

// Per-row outputs of a batch: y[r] = w3 . silu(H2[r]) + b3. Workgroup b
// takes 64-row steps b, b + G, ... (G = gridDim.x); the next step's X rows
// are read into registers while a step is evaluated.
extern "C" __global__ void __launch_bounds__(256) gd_mlp_fwd_kernel(const unsigned short *X,
                                                                    int64_t n_rows, SurW W,
                                                                    float *y) {
  __shared__ __attribute__((aligned(16))) unsigned short xs[SUR_STEP * SUR_XS];
  __shared__ __attribute__((aligned(16))) unsigned short a1s[SUR_STEP * SUR_AS];
  __shared__ float red[4][SUR_STEP];
  const int wv = threadIdx.x >> 6, l = threadIdx.x & 63, lr = l & 15, lq = l >> 4;
  ColW cw;
  bf16x8_t unused[2][4];
  load_colw(W, wv, false, cw, unused);
  const float b3 = W.b3[0];
  const int64_t n_steps = n_rows / SUR_STEP;
  uint4 xv[2];
  if (blockIdx.x < n_steps) load_x(X, (int64_t)blockIdx.x * SUR_STEP, xv);
  for (int64_t st = blockIdx.x; st < n_steps; st += gridDim.x) {
    const int64_t r0 = st * SUR_STEP;
    put_x(xv, xs, nullptr);
    if (st + gridDim.x < n_steps) load_x(X, (st + gridDim.x) * SUR_STEP, xv);
    __syncthreads();
    f32x4_t h1[4][2], h2[4][2];
    col_layer1(xs, cw, h1);
    float a1[4][2][4];
#pragma unroll
    for (int rt = 0; rt < 4; ++rt)
#pragma unroll
      for (int c = 0; c < 2; ++c)
#pragma unroll
        for (int i = 0; i < 4; ++i) a1[rt][c][i] = silu(h1[rt][c][i]);
    put_tiles(a1, wv, a1s, nullptr);
    __syncthreads();
    col_layer2(a1s, cw, h2);
    // this wave's 32 columns of w3 . silu(H2) per row, then across the waves
#pragma unroll
    for (int rt = 0; rt < 4; ++rt)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        float v = fmaf(silu(h2[rt][0][i]), cw.w3[0], silu(h2[rt][1][i]) * cw.w3[1]);
        v += __shfl_xor(v, 1);
        v += __shfl_xor(v, 2);
        v += __shfl_xor(v, 4);
        v += __shfl_xor(v, 8);
        if (lr == 0) red[wv][16 * rt + 4 * lq + i] = v;
      }
    __syncthreads();
    if (threadIdx.x < SUR_STEP) {
      const int r = threadIdx.x;
      y[r0 + r] = ((red[0][r] + red[1][r]) + (red[2][r] + red[3][r])) + b3;
    }
    __syncthreads();  // xs, a1s and red are rewritten by the next step
  }
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

// Output gradient of each row of a batch: dyr[i] = dS of row i's pose when
// that pose lies in [p0, p1), else 0 (rows of the 64-row padding that belong
// to the next poses). One thread per row.
extern "C" __global__ void gd_mlp_dyrow_kernel(const int *pose_of_row, int64_t n_rows, int p0,
                                               int p1, const float *dS, float *dyr) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_rows) return;
  const int p = pose_of_row[i];
  dyr[i] = (p >= p0 && p < p1) ? dS[p] : 0.0f;
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

// Backward of one batch fused with its weight gradients, column-split like
// gd_mlp_fwd_kernel. Workgroup b takes the contiguous 64-row steps
// [b S / G, (b + 1) S / G) (S steps, G = gridDim.x). Per step, wave w
// recomputes H1 and H2 of its 32 columns, forms
//   dy = dyr of the row (gd_mlp_dyrow_kernel),  dH2 = dy w3 * silu'(H2),
//   dA1 = dH2 W2^T (all of dH2 from LDS, W2^T slice in registers),
//   dH1 = dA1 * silu'(H1)
// and accumulates its 32 columns of
//   dW2 += A1^T dH2   (A1^T of all waves from LDS),
//   dW1 += X^T dH1    (row 56 of dW1 is the b1 gradient)
// in MFMA accumulators, with b2, w3 and b3. Only fixed-order partials go to
// HBM: p1g[b][64][128], p2g[b][128][128], pb2[b][128], pw3[b][129]
// (summed in a fixed order by gd_mlp_update_kernel or gd_mlp_gradsum_kernel,
// so training is deterministic). The next step's X rows are read into
// registers while a step is evaluated, and a step's dy rows are read at its
// start, so neither load stalls the step.
extern "C" __global__ void __launch_bounds__(256) gd_mlp_train_kernel(
    const unsigned short *X, int64_t n_rows, const float *dyr, SurW W, float *p1g, float *p2g,
    float *pb2, float *pw3) {
  extern __shared__ __attribute__((aligned(16))) unsigned char sm_train[];
  // X [64][SUR_XS] and X^T [64][SUR_TR], two of each (even and odd steps)
  unsigned short *xb = reinterpret_cast<unsigned short *>(sm_train);
  unsigned short *a1s = xb + 2 * (SUR_STEP * SUR_XS + SUR_F * SUR_TR);  // A1 [64][SUR_AS]
  unsigned short *a1t = a1s + SUR_STEP * SUR_AS;                        // A1^T [128][SUR_TR]
  unsigned short *d2s = a1t + SUR_H * SUR_TR;                           // dH2 [64][SUR_AS]
  unsigned short *d2t = d2s + SUR_STEP * SUR_AS;                        // dH2^T [128][SUR_TR]
  unsigned short *d1t = d2t + SUR_H * SUR_TR;                           // dH1^T [128][SUR_TR]
  const int wv = threadIdx.x >> 6, l = threadIdx.x & 63, lr = l & 15, lq = l >> 4;
  ColW cw;
  bf16x8_t w2r[2][4];
  load_colw(W, wv, true, cw, w2r);
  f32x4_t g2[8][2], g1[4][2];
#pragma unroll
  for (int c = 0; c < 2; ++c) {
#pragma unroll
    for (int t = 0; t < 8; ++t) g2[t][c] = f32x4_t{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int t = 0; t < 4; ++t) g1[t][c] = f32x4_t{0.f, 0.f, 0.f, 0.f};
  }
  float gw3[2] = {0.f, 0.f}, gb2[2] = {0.f, 0.f}, gb3 = 0.0f;
  const int64_t n_steps = n_rows / SUR_STEP;
  const int64_t s_beg = n_steps * blockIdx.x / gridDim.x;
  const int64_t s_end = n_steps * (blockIdx.x + 1) / gridDim.x;
  uint4 xv[2];
  if (s_beg < s_end) load_x(X, s_beg * SUR_STEP, xv);
  for (int64_t st = s_beg; st < s_end; ++st) {
    const int64_t r0 = st * SUR_STEP;
    unsigned short *xs = xb + (st & 1) * (SUR_STEP * SUR_XS + SUR_F * SUR_TR);
    unsigned short *xt = xs + SUR_STEP * SUR_XS;
    // dy of the rows 16 rt + 4 lq + i
    float4 dyv[4];
#pragma unroll
    for (int rt = 0; rt < 4; ++rt)
      dyv[rt] = *reinterpret_cast<const float4 *>(dyr + r0 + 16 * rt + 4 * lq);
    put_x(xv, xs, xt);
    if (st + 1 < s_end) load_x(X, r0 + SUR_STEP, xv);
    __syncthreads();
    f32x4_t h1[4][2], h2[4][2];
    col_layer1(xs, cw, h1);
    {
      float a1[4][2][4];
#pragma unroll
      for (int rt = 0; rt < 4; ++rt)
#pragma unroll
        for (int c = 0; c < 2; ++c)
#pragma unroll
          for (int i = 0; i < 4; ++i) a1[rt][c][i] = silu(h1[rt][c][i]);
      put_tiles(a1, wv, a1s, a1t);
    }
    __syncthreads();
    col_layer2(a1s, cw, h2);
    // dy of the rows 16 rt + 4 lq + i; w3, b2, b3 gradients
    float dy[4][4];
#pragma unroll
    for (int rt = 0; rt < 4; ++rt) {
      dy[rt][0] = dyv[rt].x;
      dy[rt][1] = dyv[rt].y;
      dy[rt][2] = dyv[rt].z;
      dy[rt][3] = dyv[rt].w;
    }
    if (wv == 0 && lr == 0)
#pragma unroll
      for (int rt = 0; rt < 4; ++rt) gb3 += (dy[rt][0] + dy[rt][1]) + (dy[rt][2] + dy[rt][3]);
    {
      float d2[4][2][4];
#pragma unroll
      for (int rt = 0; rt < 4; ++rt)
#pragma unroll
        for (int c = 0; c < 2; ++c)
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            gw3[c] = fmaf(silu(h2[rt][c][i]), dy[rt][i], gw3[c]);
            const float g = dy[rt][i] * cw.w3[c] * dsilu(h2[rt][c][i]);
            gb2[c] += g;
            d2[rt][c][i] = g;
          }
      put_tiles(d2, wv, d2s, d2t);
    }
    __syncthreads();
    // dA1 = dH2 W2^T for this wave's columns, dH1 transposed
    {
      f32x4_t da[4][2];
#pragma unroll
      for (int rt = 0; rt < 4; ++rt) da[rt][0] = da[rt][1] = f32x4_t{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int kb = 0; kb < 4; ++kb)
#pragma unroll
        for (int rt = 0; rt < 4; ++rt) {
          const bf16x8_t a = ld8(d2s + (16 * rt + lr) * SUR_AS + 32 * kb + 8 * lq);
          da[rt][0] = mfma32(a, w2r[0][kb], da[rt][0]);
          da[rt][1] = mfma32(a, w2r[1][kb], da[rt][1]);
        }
      float d1[4][2][4];
#pragma unroll
      for (int rt = 0; rt < 4; ++rt)
#pragma unroll
        for (int c = 0; c < 2; ++c)
#pragma unroll
          for (int i = 0; i < 4; ++i) d1[rt][c][i] = da[rt][c][i] * dsilu(h1[rt][c][i]);
      put_tiles(d1, wv, nullptr, d1t);
    }
    __builtin_amdgcn_s_waitcnt(0xc07f);  // this wave's dH1^T stores are in LDS
    __builtin_amdgcn_wave_barrier();
    // weight gradients over the step's 64 rows (two K blocks of 32)
#pragma unroll
    for (int kb = 0; kb < 2; ++kb) {
      const int ko = 32 * kb + 8 * lq;
      bf16x8_t b2v[2], b1v[2];
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        b2v[c] = ld8(d2t + (32 * wv + 16 * c + lr) * SUR_TR + ko);
        b1v[c] = ld8(d1t + (32 * wv + 16 * c + lr) * SUR_TR + ko);
      }
#pragma unroll
      for (int t = 0; t < 8; ++t) {
        const bf16x8_t a = ld8(a1t + (16 * t + lr) * SUR_TR + ko);
        g2[t][0] = mfma32(a, b2v[0], g2[t][0]);
        g2[t][1] = mfma32(a, b2v[1], g2[t][1]);
      }
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const bf16x8_t a = ld8(xt + (16 * t + lr) * SUR_TR + ko);
        g1[t][0] = mfma32(a, b1v[0], g1[t][0]);
        g1[t][1] = mfma32(a, b1v[1], g1[t][1]);
      }
    }
    // no barrier here: the next step writes the other X buffer, and its
    // first barrier precedes every write to the shared A1 and dH2 tiles
  }

  // partials of this workgroup: its 4 waves own disjoint columns
  float *o2 = p2g + (int64_t)blockIdx.x * SUR_H * SUR_H;
  float *o1 = p1g + (int64_t)blockIdx.x * SUR_F * SUR_H;
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    const int col = 32 * wv + 16 * c + lr;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
#pragma unroll
      for (int t = 0; t < 8; ++t) o2[(16 * t + 4 * lq + i) * SUR_H + col] = g2[t][c][i];
#pragma unroll
      for (int t = 0; t < 4; ++t) o1[(16 * t + 4 * lq + i) * SUR_H + col] = g1[t][c][i];
    }
  }
  float *w3o = pw3 + (int64_t)blockIdx.x * (SUR_H + 1);
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    float v = gw3[c];
    v += __shfl_xor(v, 16);
    v += __shfl_xor(v, 32);
    float b = gb2[c];
    b += __shfl_xor(b, 16);
    b += __shfl_xor(b, 32);
    if (lq == 0) {
      w3o[32 * wv + 16 * c + lr] = v;
      pb2[(int64_t)blockIdx.x * SUR_H + 32 * wv + 16 * c + lr] = b;
    }
  }
  gb3 += __shfl_xor(gb3, 16);
  gb3 += __shfl_xor(gb3, 32);
  if (wv == 0 && l == 0) w3o[SUR_H] = gb3;
}

// Dynamic LDS bytes of gd_mlp_train_kernel.
constexpr int SUR_TRAIN_LDS =
    (2 * (SUR_STEP * SUR_XS + SUR_F * SUR_TR) + 2 * SUR_STEP * SUR_AS + 3 * SUR_H * SUR_TR) * 2;

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
    s = sum_partials(pw3, (int)nw, SUR_H + 1, q - SP_W3);
  }
  g[q] = s;
}

// One training update: each parameter's gradient from the fixed-order sum of
// the backward's per-workgroup partials (ng of each), then Adam on it and the
// refresh of its bf16 copies. The gradient is also left in g.
extern "C" __global__ void gd_mlp_update_kernel(const float *p1, const float *p2, const float *pb2,
                                                const float *pw3, int ng, float *p, float *m,
                                                float *v, float *g, float lr, float b1, float b2,
                                                float eps, float c1, float c2,
                                                unsigned short *W1T, unsigned short *W2T,
                                                unsigned short *W2R) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= SUR_NPARAM) return;
  float gq;
  if (q < SP_W2) gq = sum_partials(p1, ng, (int64_t)SUR_F * SUR_H, q);
  else if (q < SP_B2) gq = sum_partials(p2, ng, (int64_t)SUR_H * SUR_H, q - SP_W2);
  else if (q < SP_W3) gq = sum_partials(pb2, ng, SUR_H, q - SP_B2);
  else gq = sum_partials(pw3, ng, SUR_H + 1, q - SP_W3);
  g[q] = gq;
  const float mq = b1 * m[q] + (1.0f - b1) * gq;
  const float vq = b2 * v[q] + (1.0f - b2) * gq * gq;
  m[q] = mq;
  v[q] = vq;
  const float x = p[q] - lr * (mq / c1) / (sqrtf(vq / c2) + eps);
  p[q] = x;
  if (q < SP_W2) {
    const int f = q / SUR_H, i = q % SUR_H;
    W1T[i * SUR_F + f] = f2bf(x);
  } else if (q < SP_B2) {
    const int i = (q - SP_W2) / SUR_H, j = (q - SP_W2) % SUR_H;
    W2T[j * SUR_H + i] = f2bf(x);
    W2R[i * SUR_H + j] = f2bf(x);
  }
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
