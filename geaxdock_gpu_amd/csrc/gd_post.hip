// gd_post.hip -- post-MC clustering on the device: k-means over the nine
// energy terms of accepted-step records and one medoid per cluster.
//
// The reference clusters the records of each replica into 50 groups, then
// the collected medoids into 500 (src/post_mc.C:40-101), each time with
// clusterByKmeans (src/util.C:1801-1880) over seq_kmeans
// (src/seq_kmeans.c): centers start at the first k objects of a shuffled
// list, Lloyd passes repeat while more than `threshold` of the objects change
// cluster (at most 501 passes), and each cluster reports one member. This
// file keeps that algorithm with three changes, each for reproducibility:
//   - the shuffle is seeded: the first k objects of a uniformly random
//     permutation are drawn by k steps of Fisher-Yates from Philox
//     (the reference seeds it from the clock);
//   - the medoid is the member nearest the final center, ties to the lower
//     index (the reference never updates its running distance and so keeps
//     the last member);
//   - cluster sums are exact: each term is added as a 48.16 fixed-point
//     integer (clamped to +-2^27), so the centers do not depend on the order
//     in which device threads add, and the means are rounded once from double.
//
// Layout in HBM: features SoA [9][n] (coalesced per term), membership [n],
// per group: centers [k][9] float, sums [k][9] int64, counts [k], pass count,
// change count and a done flag. A pass is two launches over all groups at
// once: assign (one workgroup per tile of 1,024 objects; centers and partial
// sums in LDS, flushed with 64-bit atomics) and update (one workgroup per
// group). The host checks the number of live groups every 8 passes; finished
// groups cost one early return per tile.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <memory>
#include <new>
#include <thread>
#include <vector>

namespace gd {

constexpr int KM_DIM = GD_MAXWEI - 1;  // vdw .. dst
constexpr int KM_TILE = 1024;
constexpr int KM_MAX_K = 512;
constexpr int KM_MAX_GROUP = 1 << 19;  // objects per group (fixed-point sum range)
constexpr float KM_CLAMP = 134217728.0f;  // 2^27

struct KmArgs {
  const float *feat;  // [KM_DIM][n_total]
  int64_t n_total;
  const int64_t *off;  // [n_groups + 1]
  const int *tile_g;   // tile -> group
  const int *tile_s;   // tile -> first object (group-local)
  int *memb;           // [n_total], group-local cluster id or -1
  float *cent;         // [n_groups][k][KM_DIM]
  long long *sum;      // [n_groups][k][KM_DIM]
  int *cnt;            // [n_groups][k]
  int *delta, *done, *pass;  // [n_groups]
  int *active;
  unsigned long long *best;  // [n_groups][k]: (distance bits << 32) | member
  int *size;                 // [n_groups][k]
  int k;
  float threshold;
  int max_pass;
};

__host__ __device__ inline long long km_fixed(float v) {
  if (!(v == v)) return 0;
  v = v > KM_CLAMP ? KM_CLAMP : (v < -KM_CLAMP ? -KM_CLAMP : v);
  return (long long)rint((double)v * 65536.0);
}

// Squared Euclidean distance: terms in order, each square added with one
// fused multiply-add (the CPU restatement calls fmaf the same way).
__host__ __device__ inline float km_dist2(const float *a, const float *b) {
#pragma clang fp contract(off)
  float s = 0.0f;
#pragma unroll
  for (int d = 0; d < KM_DIM; ++d) {
    const float t = a[d] - b[d];
    s = fmaf(t, t, s);
  }
  return s;
}

// The centers are read-only here and wave-uniform: passed as their own
// restrict pointer they are fetched with scalar loads and feed the distance
// arithmetic as scalar operands (no LDS traffic for them).
__global__ void __launch_bounds__(256) gd_km_assign_kernel(KmArgs A,
                                                           const float *__restrict__ cent) {
  const int g = A.tile_g[blockIdx.x];
  if (A.done[g]) return;
  const int k = A.k;
  extern __shared__ __align__(16) unsigned char km_lds[];
  long long *sum = (long long *)km_lds;  // [k][KM_DIM]
  int *cnt = (int *)(sum + (size_t)k * KM_DIM);  // [k]
  int *dlt = cnt + k;
  const float *cen = cent + (size_t)g * k * KM_DIM;
  for (int j = threadIdx.x; j < k * KM_DIM; j += blockDim.x) sum[j] = 0;
  for (int c = threadIdx.x; c < k; c += blockDim.x) cnt[c] = 0;
  if (threadIdx.x == 0) *dlt = 0;
  __syncthreads();
  const int64_t base = A.off[g];
  const int n = (int)(A.off[g + 1] - base);
  const int s0 = A.tile_s[blockIdx.x];
  const int s1 = min(n, s0 + KM_TILE);
  int changed = 0;
  // two objects per lane share each center's scalar loads
  const int stride = 2 * (int)blockDim.x;
  for (int i0 = s0 + (int)threadIdx.x; i0 < s1; i0 += stride) {
    const int i1 = i0 + (int)blockDim.x;
    const bool has1 = i1 < s1;
    float f0[KM_DIM], f1[KM_DIM];
#pragma unroll
    for (int d = 0; d < KM_DIM; ++d) {
      f0[d] = A.feat[(size_t)d * A.n_total + base + i0];
      f1[d] = has1 ? A.feat[(size_t)d * A.n_total + base + i1] : f0[d];
    }
    int b0 = 0, b1 = 0;
    float d0 = km_dist2(f0, cen), d1 = km_dist2(f1, cen);
#pragma unroll 2
    for (int c = 1; c < k; ++c) {
      const float *cc = cen + (size_t)c * KM_DIM;
      const float e0 = km_dist2(f0, cc), e1 = km_dist2(f1, cc);
      if (e0 < d0) {
        d0 = e0;
        b0 = c;
      }
      if (e1 < d1) {
        d1 = e1;
        b1 = c;
      }
    }
    for (int h = 0; h < 2; ++h) {
      if (h == 1 && !has1) break;
      const int i = h ? i1 : i0, bc = h ? b1 : b0;
      const float *f = h ? f1 : f0;
      int *m = A.memb + base + i;
      if (*m != bc) {
        changed += 1;
        *m = bc;
      }
#pragma unroll
      for (int d = 0; d < KM_DIM; ++d)
        atomicAdd((unsigned long long *)&sum[bc * KM_DIM + d], (unsigned long long)km_fixed(f[d]));
      atomicAdd(&cnt[bc], 1);
    }
  }
  if (changed) atomicAdd(dlt, changed);
  __syncthreads();
  long long *gs = A.sum + (size_t)g * k * KM_DIM;
  int *gn = A.cnt + (size_t)g * k;
  for (int c = threadIdx.x; c < k; c += blockDim.x) {
    if (!cnt[c]) continue;
    atomicAdd(gn + c, cnt[c]);
    for (int d = 0; d < KM_DIM; ++d)
      if (sum[c * KM_DIM + d])
        atomicAdd((unsigned long long *)&gs[c * KM_DIM + d], (unsigned long long)sum[c * KM_DIM + d]);
  }
  if (threadIdx.x == 0 && *dlt) atomicAdd(A.delta + g, *dlt);
}

__host__ __device__ inline float km_mean(long long s, int n) {
  return (float)((double)s / (double)n / 65536.0);
}

__global__ void __launch_bounds__(256) gd_km_update_kernel(KmArgs A) {
  const int g = blockIdx.x;
  if (A.done[g]) return;
  const int k = A.k;
  float *gc = A.cent + (size_t)g * k * KM_DIM;
  long long *gs = A.sum + (size_t)g * k * KM_DIM;
  int *gn = A.cnt + (size_t)g * k;
  for (int j = threadIdx.x; j < k * KM_DIM; j += blockDim.x) {
    const int n = gn[j / KM_DIM];
    if (n > 0) gc[j] = km_mean(gs[j], n);
    gs[j] = 0;
  }
  __syncthreads();
  for (int c = threadIdx.x; c < k; c += blockDim.x) gn[c] = 0;
  if (threadIdx.x == 0) {
    const int n = (int)(A.off[g + 1] - A.off[g]);
    const float frac = (float)A.delta[g] / (float)n;
    A.delta[g] = 0;
    const int p = A.pass[g] + 1;
    A.pass[g] = p;
    if (!(frac > A.threshold) || p >= A.max_pass) A.done[g] = 1;
    else atomicAdd(A.active, 1);
  }
}

__global__ void __launch_bounds__(256) gd_km_medoid_kernel(KmArgs A) {
  const int g = A.tile_g[blockIdx.x];
  const int k = A.k;
  extern __shared__ __align__(16) unsigned char km_lds[];
  float *cen = (float *)km_lds;
  const float *gc = A.cent + (size_t)g * k * KM_DIM;
  for (int j = threadIdx.x; j < k * KM_DIM; j += blockDim.x) cen[j] = gc[j];
  __syncthreads();
  const int64_t base = A.off[g];
  const int n = (int)(A.off[g + 1] - base);
  const int s0 = A.tile_s[blockIdx.x];
  const int s1 = min(n, s0 + KM_TILE);
  for (int i = s0 + (int)threadIdx.x; i < s1; i += blockDim.x) {
    const int c = A.memb[base + i];
    if (c < 0 || c >= k) continue;
    float f[KM_DIM];
#pragma unroll
    for (int d = 0; d < KM_DIM; ++d) f[d] = A.feat[(size_t)d * A.n_total + base + i];
    const float dd = km_dist2(f, cen + (size_t)c * KM_DIM);
    const unsigned long long key = ((unsigned long long)__float_as_uint(dd) << 32) | (unsigned)i;
    atomicMin(A.best + (size_t)g * k + c, key);
    atomicAdd(A.size + (size_t)g * k + c, 1);
  }
}

// First k positions of a uniformly random permutation of [0, n): k steps of
// Fisher-Yates, position i swapped with i + floor(u * (n - i)), u from
// Philox (i, group, level, 0x6B6D) under the seed.
inline void km_init_indices(int n, int k, int group, int level, uint64_t seed, std::vector<int> &perm,
                            int *out) {
  if ((int)perm.size() < n) {
    const int old = (int)perm.size();
    perm.resize(n);
    for (int i = old; i < n; ++i) perm[i] = i;
  }
  std::vector<int> touched;
  for (int i = 0; i < k; ++i) {
    const u32x4 r = philox4x32_10({(uint32_t)i, (uint32_t)group, (uint32_t)level, 0x6B6Du},
                                  (uint32_t)seed, (uint32_t)(seed >> 32));
    const int j = i + (int)(((uint64_t)r.x * (uint64_t)(n - i)) >> 32);
    std::swap(perm[i], perm[j]);
    out[i] = perm[i];
    touched.push_back(i);
    touched.push_back(j);
  }
  for (int t : touched) perm[t] = t;  // back to the identity for the next group
}

}  // namespace gd

namespace {

// Host loops over millions of records run on up to 16 threads.
template <class F>
void par_for(int64_t n, F f) {
  const int64_t hw = std::max(1u, std::thread::hardware_concurrency());
  const int nt = (int)std::min<int64_t>(std::min<int64_t>(hw, 16), std::max<int64_t>(1, n / 65536));
  if (nt <= 1) {
    f(0, (int64_t)0, n);
    return;
  }
  std::vector<std::thread> th;
  for (int t = 0; t < nt; ++t)
    th.emplace_back([&, t] { f(t, n * t / nt, n * (t + 1) / nt); });
  for (auto &x : th) x.join();
}
// Host arrays that every element of is written before use: no zero fill.
template <class T>
std::unique_ptr<T[]> uninit(size_t n) {
  return std::unique_ptr<T[]>(new T[std::max<size_t>(n, 1)]);
}

inline int par_threads(int64_t n) {
  const int64_t hw = std::max(1u, std::thread::hardware_concurrency());
  return (int)std::min<int64_t>(std::min<int64_t>(hw, 16), std::max<int64_t>(1, n / 65536));
}

// Device k-means of n_groups groups of objects whose features are given as
// SoA [9][n_total] on the host. medoid/size: [n_groups][k], group-local
// member index or -1.
gd_status km_run(gd_ctx *ctx, int n_groups, const int64_t *off, const float *soa_in, int k,
                 float threshold, int max_pass, uint64_t seed, int level, int *medoid, int *size,
                 int *passes, double *device_ms) {
  using namespace gd;
  if (n_groups < 0 || !off || (!soa_in && off[n_groups] > 0) || !medoid || !size)
    return fail(ctx, GD_ERR_INVALID, "kmeans: bad arguments");
  if (k < 1 || k > KM_MAX_K) return fail(ctx, GD_ERR_CAPACITY, "kmeans: k %d outside [1, %d]", k, KM_MAX_K);
  if (max_pass < 1) return fail(ctx, GD_ERR_INVALID, "kmeans: max_pass %d < 1", max_pass);
  if (off[0] != 0) return fail(ctx, GD_ERR_INVALID, "kmeans: group offsets must start at 0");
  for (int g = 0; g < n_groups; ++g) {
    const int64_t n = off[g + 1] - off[g];
    if (n < 0) return fail(ctx, GD_ERR_INVALID, "kmeans: group %d has negative size", g);
    if (n > KM_MAX_GROUP)
      return fail(ctx, GD_ERR_CAPACITY, "kmeans: group %d has %lld objects (max %d)", g,
                  (long long)n, KM_MAX_GROUP);
  }
  const int64_t n_total = n_groups ? off[n_groups] : 0;
  if (n_total > INT32_MAX) return fail(ctx, GD_ERR_CAPACITY, "kmeans: %lld objects", (long long)n_total);
  for (int i = 0; i < n_groups * k; ++i) {
    medoid[i] = -1;
    size[i] = 0;
  }
  // groups with fewer than k objects: every object is its own medoid
  std::vector<int> tile_g, tile_s;
  std::vector<float> cent((size_t)n_groups * k * KM_DIM, 0.0f);
  std::vector<int> done(n_groups, 1), init(k);
  std::vector<int> perm;
  for (int g = 0; g < n_groups; ++g) {
    const int n = (int)(off[g + 1] - off[g]);
    if (passes) passes[g] = 0;
    if (n < k) {
      for (int i = 0; i < n; ++i) {
        medoid[(size_t)g * k + i] = i;
        size[(size_t)g * k + i] = 1;
      }
      continue;
    }
    done[g] = 0;
    km_init_indices(n, k, g, level, seed, perm, init.data());
    for (int c = 0; c < k; ++c)
      for (int d = 0; d < KM_DIM; ++d)
        cent[((size_t)g * k + c) * KM_DIM + d] = soa_in[(size_t)d * n_total + off[g] + init[c]];
    for (int s = 0; s < n; s += KM_TILE) {
      tile_g.push_back(g);
      tile_s.push_back(s);
    }
  }
  const int n_tiles = (int)tile_g.size();
  if (n_tiles == 0) return GD_OK;
  HIPCHK(hipSetDevice(ctx->device));
  const size_t gk = (size_t)n_groups * k;
  std::vector<void *> bufs;
  auto alloc = [&](void **p, size_t bytes) -> hipError_t {
    const hipError_t e = hipMalloc(p, std::max<size_t>(bytes, 16));
    if (e == hipSuccess) bufs.push_back(*p);
    return e;
  };
  struct Free {
    std::vector<void *> &b;
    ~Free() {
      for (void *p : b) (void)hipFree(p);
    }
  } free_all{bufs};
  KmArgs A{};
  float *d_feat;
  int64_t *d_off;
  int *d_tg, *d_ts;
  HIPCHK(alloc((void **)&d_feat, (size_t)KM_DIM * n_total * 4));
  HIPCHK(alloc((void **)&d_off, (size_t)(n_groups + 1) * 8));
  HIPCHK(alloc((void **)&d_tg, (size_t)n_tiles * 4));
  HIPCHK(alloc((void **)&d_ts, (size_t)n_tiles * 4));
  HIPCHK(alloc((void **)&A.memb, (size_t)n_total * 4));
  HIPCHK(alloc((void **)&A.cent, gk * KM_DIM * 4));
  HIPCHK(alloc((void **)&A.sum, gk * KM_DIM * 8));
  HIPCHK(alloc((void **)&A.cnt, gk * 4));
  HIPCHK(alloc((void **)&A.delta, (size_t)n_groups * 4));
  HIPCHK(alloc((void **)&A.done, (size_t)n_groups * 4));
  HIPCHK(alloc((void **)&A.pass, (size_t)n_groups * 4));
  HIPCHK(alloc((void **)&A.active, 4));
  HIPCHK(alloc((void **)&A.best, gk * 8));
  HIPCHK(alloc((void **)&A.size, gk * 4));
  hipStream_t st = ctx->stream;
  HIPCHK(hipMemcpyAsync(d_feat, soa_in, (size_t)KM_DIM * n_total * 4, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(d_off, off, (size_t)(n_groups + 1) * 8, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(d_tg, tile_g.data(), (size_t)n_tiles * 4, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(d_ts, tile_s.data(), (size_t)n_tiles * 4, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(A.cent, cent.data(), gk * KM_DIM * 4, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(A.done, done.data(), (size_t)n_groups * 4, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemsetAsync(A.memb, 0xFF, (size_t)n_total * 4, st));
  HIPCHK(hipMemsetAsync(A.sum, 0, gk * KM_DIM * 8, st));
  HIPCHK(hipMemsetAsync(A.cnt, 0, gk * 4, st));
  HIPCHK(hipMemsetAsync(A.delta, 0, (size_t)n_groups * 4, st));
  HIPCHK(hipMemsetAsync(A.pass, 0, (size_t)n_groups * 4, st));
  HIPCHK(hipMemsetAsync(A.best, 0xFF, gk * 8, st));
  HIPCHK(hipMemsetAsync(A.size, 0, gk * 4, st));
  A.feat = d_feat;
  A.n_total = n_total;
  A.off = d_off;
  A.tile_g = d_tg;
  A.tile_s = d_ts;
  A.k = k;
  A.threshold = threshold;
  A.max_pass = max_pass;
  const size_t lds_assign = (size_t)k * KM_DIM * 8 + (size_t)k * 4 + 16;
  const size_t lds_medoid = (size_t)k * KM_DIM * 4;
  HIPCHK(hipFuncSetAttribute((const void *)gd_km_assign_kernel,
                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_assign));
  HIPCHK(hipEventRecord(ctx->ev0, st));
  for (int p = 0; p < max_pass;) {
    const int batch = std::min(8, max_pass - p);
    for (int b = 0; b < batch; ++b) {
      if (b == batch - 1) HIPCHK(hipMemsetAsync(A.active, 0, 4, st));
      hipLaunchKernelGGL(gd_km_assign_kernel, dim3(n_tiles), dim3(256), lds_assign, st, A,
                         (const float *)A.cent);
      HIPCHK(hipGetLastError());
      hipLaunchKernelGGL(gd_km_update_kernel, dim3(n_groups), dim3(256), 0, st, A);
      HIPCHK(hipGetLastError());
    }
    p += batch;
    int live = 0;
    HIPCHK(hipMemcpyAsync(&live, A.active, 4, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    if (live == 0) break;
  }
  hipLaunchKernelGGL(gd_km_medoid_kernel, dim3(n_tiles), dim3(256), lds_medoid, st, A);
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventRecord(ctx->ev1, st));
  std::vector<unsigned long long> best(gk);
  std::vector<int> sz(gk), pass(n_groups);
  HIPCHK(hipMemcpyAsync(best.data(), A.best, gk * 8, hipMemcpyDeviceToHost, st));
  HIPCHK(hipMemcpyAsync(sz.data(), A.size, gk * 4, hipMemcpyDeviceToHost, st));
  HIPCHK(hipMemcpyAsync(pass.data(), A.pass, (size_t)n_groups * 4, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  if (device_ms) {
    float ms = 0.0f;
    HIPCHK(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    *device_ms += ms;
  }
  for (int g = 0; g < n_groups; ++g) {
    if (off[g + 1] - off[g] < k) continue;
    if (passes) passes[g] = pass[g];
    for (int c = 0; c < k; ++c) {
      const size_t j = (size_t)g * k + c;
      if (sz[j] > 0) {
        medoid[j] = (int)(best[j] & 0xFFFFFFFFull);
        size[j] = sz[j];
      }
    }
  }
  return GD_OK;
}

}  // namespace

extern "C" {

gd_status gd_kmeans_medoids(gd_ctx *ctx, int n_groups, const int64_t *group_off,
                            const float *features, int k, float threshold, int max_pass,
                            uint64_t seed, int *medoid, int *size, int *passes,
                            double *device_ms) {
  if (!ctx) return GD_ERR_INVALID;
  if (device_ms) *device_ms = 0.0;
  try {
    using namespace gd;
    if (n_groups < 0 || !group_off || (n_groups > 0 && group_off[n_groups] > INT32_MAX))
      return fail(ctx, GD_ERR_INVALID, "kmeans: bad group offsets");
    const int64_t n_total = n_groups > 0 ? std::max<int64_t>(group_off[n_groups], 0) : 0;
    if (n_total > 0 && !features) return fail(ctx, GD_ERR_INVALID, "kmeans: null features");
    // row-major objects to SoA terms
    auto soa_v = uninit<float>((size_t)KM_DIM * n_total);
    float *soa = soa_v.get();
    par_for(n_total, [&](int, int64_t a, int64_t b) {
      for (int d = 0; d < KM_DIM; ++d)
        for (int64_t i = a; i < b; ++i) soa[(size_t)d * n_total + i] = features[i * KM_DIM + d];
    });
    return km_run(ctx, n_groups, group_off, soa, k, threshold, max_pass, seed, 0, medoid, size,
                  passes, device_ms);
  } catch (const std::bad_alloc &) {
    return fail(ctx, GD_ERR_NOMEM, "kmeans: out of host memory");
  }
}

}  // extern "C"

namespace {
gd_status post_mc_impl(gd_ctx *ctx, const gd_record *recs, int64_t n, int k_replica, int k_final,
                       uint64_t seed, int64_t *medoid_rec, int *medoid_size, int *n_medoids,
                       double *device_ms);
}  // namespace

extern "C" {

gd_status gd_post_mc(gd_ctx *ctx, const gd_record *recs, int64_t n, int k_replica, int k_final,
                     uint64_t seed, int64_t *medoid_rec, int *medoid_size, int *n_medoids,
                     double *device_ms) {
  if (!ctx) return GD_ERR_INVALID;
  try {
    return post_mc_impl(ctx, recs, n, k_replica, k_final, seed, medoid_rec, medoid_size,
                        n_medoids, device_ms);
  } catch (const std::bad_alloc &) {
    return fail(ctx, GD_ERR_NOMEM, "post_mc: out of host memory");
  }
}

}  // extern "C"

namespace {
gd_status post_mc_impl(gd_ctx *ctx, const gd_record *recs, int64_t n, int k_replica, int k_final,
                       uint64_t seed, int64_t *medoid_rec, int *medoid_size, int *n_medoids,
                       double *device_ms) {
  using namespace gd;
  if (device_ms) *device_ms = 0.0;
  if (n < 0 || (n > 0 && !recs) || !medoid_rec || !medoid_size || !n_medoids)
    return fail(ctx, GD_ERR_INVALID, "post_mc: bad arguments");
  if (k_replica < 1 || k_replica > KM_MAX_K || k_final < 1 || k_final > KM_MAX_K)
    return fail(ctx, GD_ERR_CAPACITY, "post_mc: k %d / %d outside [1, %d]", k_replica, k_final,
                KM_MAX_K);
  *n_medoids = 0;
  if (n == 0) return GD_OK;
  // records grouped by replica (ascending replica, each in the given order):
  // a stable counting sort, chunk-parallel
  const int nt = par_threads(n);
  std::vector<int> tmax(nt, -1), tbad(nt, -1);
  par_for(n, [&](int t, int64_t a, int64_t b) {
    int mx = -1;
    for (int64_t i = a; i < b; ++i) {
      const int r = recs[i].replica.idx_rep;
      if (r < 0 && tbad[t] < 0) tbad[t] = (int)std::min<int64_t>(i, INT32_MAX);
      mx = std::max(mx, r);
    }
    tmax[t] = mx;
  });
  for (int t = 0; t < nt; ++t)
    if (tbad[t] >= 0)
      return fail(ctx, GD_ERR_INVALID, "post_mc: record %d has a negative replica", tbad[t]);
  const int max_rep = *std::max_element(tmax.begin(), tmax.end());
  const size_t nr = (size_t)max_rep + 1;
  // per-thread counts, then each thread's first slot per replica
  std::vector<int64_t> tcnt((size_t)nt * nr, 0);
  par_for(n, [&](int t, int64_t a, int64_t b) {
    int64_t *c = tcnt.data() + (size_t)t * nr;
    for (int64_t i = a; i < b; ++i) c[recs[i].replica.idx_rep] += 1;
  });
  std::vector<int64_t> cnt(nr + 1, 0);
  for (size_t r = 0; r < nr; ++r) {
    int64_t s0 = cnt[r];
    for (int t = 0; t < nt; ++t) {
      const int64_t c = tcnt[(size_t)t * nr + r];
      tcnt[(size_t)t * nr + r] = s0;
      s0 += c;
    }
    cnt[r + 1] = s0;
  }
  auto order = uninit<int64_t>((size_t)n);
  par_for(n, [&](int t, int64_t a, int64_t b) {
    int64_t *pos = tcnt.data() + (size_t)t * nr;
    for (int64_t i = a; i < b; ++i) order[(size_t)pos[recs[i].replica.idx_rep]++] = i;
  });
  // groups: replicas that have records
  std::vector<int64_t> off{0};
  for (size_t r = 0; r < nr; ++r)
    if (cnt[r + 1] > cnt[r]) off.push_back(cnt[r + 1]);
  const int n_groups = (int)off.size() - 1;
  // features of the grouped records as SoA terms
  auto feat_v = uninit<float>((size_t)KM_DIM * n);
  float *feat = feat_v.get();
  par_for(n, [&](int, int64_t a, int64_t b) {
    for (int64_t i = a; i < b; ++i) {
      const float *e = recs[order[i]].energy.e;
      for (int d = 0; d < KM_DIM; ++d) feat[(size_t)d * n + i] = e[d];
    }
  });
  std::vector<int> med((size_t)n_groups * k_replica), sz((size_t)n_groups * k_replica);
  TRY(km_run(ctx, n_groups, off.data(), feat, k_replica, 0.001f, 501, seed, 1, med.data(),
             sz.data(), nullptr, device_ms));
  // level 2 over the level-1 medoids, in replica and cluster order
  std::vector<int64_t> l1;
  for (int g = 0; g < n_groups; ++g)
    for (int c = 0; c < k_replica; ++c) {
      const int m = med[(size_t)g * k_replica + c];
      if (m >= 0) l1.push_back(order[off[g] + m]);
    }
  const size_t n1 = l1.size();
  std::vector<float> f2((size_t)KM_DIM * std::max<size_t>(n1, 1));
  for (size_t i = 0; i < n1; ++i)
    for (int d = 0; d < KM_DIM; ++d) f2[(size_t)d * n1 + i] = recs[l1[i]].energy.e[d];
  const int64_t off2[2] = {0, (int64_t)l1.size()};
  std::vector<int> med2(k_final), sz2(k_final);
  TRY(km_run(ctx, 1, off2, f2.data(), k_final, 0.001f, 501, seed, 2, med2.data(), sz2.data(),
             nullptr, device_ms));
  int m = 0;
  for (int c = 0; c < k_final; ++c)
    if (med2[c] >= 0) {
      medoid_rec[m] = l1[med2[c]];
      medoid_size[m] = sz2[c];
      m += 1;
    }
  *n_medoids = m;
  return GD_OK;
}
}  // namespace
