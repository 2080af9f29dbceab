/*
 * post.c -- CPU oracle of post-MC clustering. TEST INFRASTRUCTURE ONLY:
 * called by tests/, never by the engine.
 *
 * Follows clusterByKmeans (src/util.C:1801-1880) over seq_kmeans
 * (src/seq_kmeans.c: find_nearest_cluster, the pass loop with its
 * "delta > threshold && loop++ < 500" test) and post_mc
 * (src/post_mc.C:40-101), one object at a time in object order, with the
 * engine's three stated changes: a seeded choice of the initial centers
 * (k Fisher-Yates steps from Philox), the medoid as the member nearest the
 * final center, and exact 48.16 fixed-point cluster sums.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "gdo.h"

#define KM_DIM (GD_MAXWEI - 1)

static long long fixed16(float v) {
  const float lim = 134217728.0f; /* 2^27 */
  if (!(v == v)) return 0;
  if (v > lim) v = lim;
  if (v < -lim) v = -lim;
  return (long long)rint((double)v * 65536.0);
}

static float dist2(const float *a, const float *b) {
  float s = 0.0f;
  for (int d = 0; d < KM_DIM; ++d) {
    const float t = a[d] - b[d];
    s = fmaf(t, t, s);
  }
  return s;
}

static uint32_t fbits(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  return u;
}

void gdo_km_init(int n, int k, int group, int level, uint64_t seed, int *out) {
  int *perm = (int *)malloc(sizeof(int) * (size_t)n);
  for (int i = 0; i < n; ++i) perm[i] = i;
  const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  for (int i = 0; i < k; ++i) {
    const uint32_t ctr[4] = {(uint32_t)i, (uint32_t)group, (uint32_t)level, 0x6B6Du};
    uint32_t r[4];
    gdo_philox4x32_10(ctr, key, r);
    const int j = i + (int)(((uint64_t)r[0] * (uint64_t)(n - i)) >> 32);
    const int t = perm[i];
    perm[i] = perm[j];
    perm[j] = t;
    out[i] = perm[i];
  }
  free(perm);
}

/* One group: f[n][9]. memb[n], medoid[k], size[k]; returns the pass count. */
int gdo_kmeans(const float *f, int n, int k, float threshold, int max_pass, int group, int level,
               uint64_t seed, int *memb, int *medoid, int *size) {
  if (k < 1 || n < 0) return 0;
  for (int c = 0; c < k; ++c) {
    medoid[c] = -1;
    size[c] = 0;
  }
  if (n < k) {
    for (int i = 0; i < n; ++i) {
      medoid[i] = i;
      size[i] = 1;
      memb[i] = i;
    }
    return 0;
  }
  float *cen = (float *)malloc(sizeof(float) * (size_t)k * KM_DIM);
  long long *sum = (long long *)calloc((size_t)k * KM_DIM, sizeof(long long));
  int *cnt = (int *)calloc((size_t)k, sizeof(int));
  int *init = (int *)malloc(sizeof(int) * (size_t)k);
  gdo_km_init(n, k, group, level, seed, init);
  for (int c = 0; c < k; ++c)
    for (int d = 0; d < KM_DIM; ++d) cen[c * KM_DIM + d] = f[(size_t)init[c] * KM_DIM + d];
  for (int i = 0; i < n; ++i) memb[i] = -1;
  int pass = 0;
  for (;;) {
    int changed = 0;
    for (int i = 0; i < n; ++i) {
      const float *o = f + (size_t)i * KM_DIM;
      int bc = 0;
      float bd = dist2(o, cen);
      for (int c = 1; c < k; ++c) {
        const float dd = dist2(o, cen + (size_t)c * KM_DIM);
        if (dd < bd) {
          bd = dd;
          bc = c;
        }
      }
      if (memb[i] != bc) changed += 1;
      memb[i] = bc;
      cnt[bc] += 1;
      for (int d = 0; d < KM_DIM; ++d) sum[bc * KM_DIM + d] += fixed16(o[d]);
    }
    for (int c = 0; c < k; ++c) {
      for (int d = 0; d < KM_DIM; ++d) {
        if (cnt[c] > 0)
          cen[c * KM_DIM + d] = (float)((double)sum[c * KM_DIM + d] / (double)cnt[c] / 65536.0);
        sum[c * KM_DIM + d] = 0;
      }
      cnt[c] = 0;
    }
    pass += 1;
    const float frac = (float)changed / (float)n;
    if (!(frac > threshold) || pass >= max_pass) break;
  }
  /* medoid: smallest (distance bits, index) among the members */
  uint64_t *best = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)k);
  for (int c = 0; c < k; ++c) best[c] = ~(uint64_t)0;
  for (int i = 0; i < n; ++i) {
    const int c = memb[i];
    const uint64_t key = ((uint64_t)fbits(dist2(f + (size_t)i * KM_DIM, cen + (size_t)c * KM_DIM))
                          << 32) | (uint32_t)i;
    if (key < best[c]) best[c] = key;
    size[c] += 1;
  }
  for (int c = 0; c < k; ++c)
    if (size[c] > 0) medoid[c] = (int)(best[c] & 0xFFFFFFFFu);
  free(best);
  free(cen);
  free(sum);
  free(cnt);
  free(init);
  return pass;
}

/* post_mc: records grouped by replica (ascending id, given order), k_rep
 * clusters each, then k_final over the medoids in replica and cluster
 * order. Returns the medoid count; medoid_rec/medoid_size hold k_final. */
int gdo_post_mc(const gd_record *recs, int64_t n, int k_rep, int k_final, uint64_t seed,
                int64_t *medoid_rec, int *medoid_size) {
  int max_rep = -1;
  for (int64_t i = 0; i < n; ++i)
    if (recs[i].replica.idx_rep > max_rep) max_rep = recs[i].replica.idx_rep;
  int64_t *l1 = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
  int64_t n1 = 0;
  int64_t *idx = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
  float *f = (float *)malloc(sizeof(float) * KM_DIM * (size_t)(n > 0 ? n : 1));
  int *memb = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
  int *med = (int *)malloc(sizeof(int) * (size_t)(k_rep > k_final ? k_rep : k_final));
  int *sz = (int *)malloc(sizeof(int) * (size_t)(k_rep > k_final ? k_rep : k_final));
  int group = 0;
  for (int r = 0; r <= max_rep; ++r) {
    int m = 0;
    for (int64_t i = 0; i < n; ++i)
      if (recs[i].replica.idx_rep == r) {
        idx[m] = i;
        memcpy(f + (size_t)m * KM_DIM, recs[i].energy.e, sizeof(float) * KM_DIM);
        m += 1;
      }
    if (m == 0) continue;
    gdo_kmeans(f, m, k_rep, 0.001f, 501, group, 1, seed, memb, med, sz);
    group += 1;
    for (int c = 0; c < k_rep; ++c)
      if (med[c] >= 0) l1[n1++] = idx[med[c]];
  }
  for (int64_t i = 0; i < n1; ++i)
    memcpy(f + (size_t)i * KM_DIM, recs[l1[i]].energy.e, sizeof(float) * KM_DIM);
  gdo_kmeans(f, (int)n1, k_final, 0.001f, 501, 0, 2, seed, memb, med, sz);
  int out = 0;
  for (int c = 0; c < k_final; ++c)
    if (med[c] >= 0) {
      medoid_rec[out] = l1[med[c]];
      medoid_size[out] = sz[c];
      out += 1;
    }
  free(l1);
  free(idx);
  free(f);
  free(memb);
  free(med);
  free(sz);
  return out;
}
