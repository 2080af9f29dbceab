// gd_mc.h -- the MC step kernel for CDNA4 (gfx950): device code and the
// gd_mc_kernel template, instantiated per atoms-per-lane in gd_mc_inst.hip.
//
// One workgroup owns a pack: whole temperature ladders of one complex that
// share one protein conformer. The receptor conformer, the type-pair
// parameter records and the KDE points are staged into LDS once per launch;
// each replica is `lpr` lanes of one wave (8..64), each lane owns `APL`
// ligand atoms, and every replica-level sum is a butterfly over its lanes, so
// the MC step needs no workgroup barrier. Barriers occur only at replica-
// exchange points, where one thread per ladder performs the swap sweep.
//
// Reference step order (src/kernel_cuda_l1_montecarlo.cu:95-119): Move_d,
// CalcRmsd_d, CalcMcc_d, CalcEnergy_d (+CombineEnergy_d), Accept_d,
// RecordLigand_d; ExchangeReplicas_d between segments
// (src/kernel_cuda_lancher_siglegpu.cu:25-31).
#pragma once

#include <type_traits>

#include "gd_device.h"
#include "gd_kernels.h"
#include "gd_sur.h"

namespace gd {

struct Lds {
  const float4 *xyze;
  const float *hpp;
  const int *psp;
  const float *par;
  const float4 *kde;  // LDS, or HBM for tiled receptors
  const int2 *kr;
  const int4 *ch;
  float *xe;
  int *xi;
  int *xa;
  int *xo;
  float *xp;
  // tiled receptors: the pack's conformer in HBM (sorted order) and the two
  // LDS tile buffers of `tile` points each
  const float4 *gx;
  const float *ghpp;
  const int *gpsp;
  float4 *tx;
  float *th;
  int *tp;
  int tile;
};

// Receptor split of one workgroup: lanes [h*half, (h+1)*half) form group h
// of ns (1, 2, 4 or 8, per pack); every group holds the same replicas and
// covers its share of each receptor chunk. `par` selects the LDS parity
// buffer of the next exchange (each exchange flips it, so one barrier per
// exchange suffices).
struct Split {
  int h, ns, half, ltid, slot, slots;
  int xp_floats;
  int par;
};

__device__ __forceinline__ float seg_sum(float v, int lpr) {
  for (int m = 1; m < lpr; m <<= 1) v += __shfl_xor(v, m);
  return v;
}
__device__ __forceinline__ int seg_sum_i(int v, int lpr) {
  for (int m = 1; m < lpr; m <<= 1) v += __shfl_xor(v, m);
  return v;
}

// KDE sum of one atom over points [b, e) of its type's range:
// sum_j exp2(kde2l |x - q_j|^2) (src/kernel_cuda_l2_calcenergy.cu:232-246).
// Four points are read and exponentiated ahead of their additions, which stay
// in point order, so the sum has the bits of the one-point-at-a-time loop
// while the LDS read latency of the next points overlaps the arithmetic.
__device__ __forceinline__ float kde_sum(const float4 *kde, float kde2l, float x, float y,
                                         float z, int b, int e) {
  float ks = 0.0f;
  int j = b;
  for (; j + 4 <= e; j += 4) {
    const float4 Q0 = kde[j], Q1 = kde[j + 1], Q2 = kde[j + 2], Q3 = kde[j + 3];
    const float e0 = fast_exp2(kde2l * pair_d2(x, y, z, Q0.x, Q0.y, Q0.z));
    const float e1 = fast_exp2(kde2l * pair_d2(x, y, z, Q1.x, Q1.y, Q1.z));
    const float e2 = fast_exp2(kde2l * pair_d2(x, y, z, Q2.x, Q2.y, Q2.z));
    const float e3 = fast_exp2(kde2l * pair_d2(x, y, z, Q3.x, Q3.y, Q3.z));
    ks += e0;
    ks += e1;
    ks += e2;
    ks += e3;
  }
  for (; j < e; ++j) {
    const float4 Q = kde[j];
    ks += fast_exp2(kde2l * pair_d2(x, y, z, Q.x, Q.y, Q.z));
  }
  return ks;
}

// HPC switching polynomial (src/kernel_cuda_l2_calcenergy.cu:116-121):
// 1 - (3.5 x - 4.5 x^2 + 2.5 x^3 - 0.5 x^4), x = d^2 / 81, which factors as
// (1 - x)^3 (1 - x/2) = u^3 (1 + u) / 2 with u = 1 - x. Clamping u at 0 makes
// it exactly 0 beyond d = 9 (the reference's d^2 <= 81 test; u never exceeds
// 1, so the clamp to [0, 1] of the fma's output modifier does it), and the factored
// form has no cancellation near the cut-off. The 1/2 is folded into the
// receptor's hpp values when they are staged into LDS (an exact scaling).
constexpr float INV81 = 1.0f / 81.0f;
constexpr float LOG2E = 1.44269504088896340736f;

struct PairConst {
  float el0, el1, a1, b1;
};

// Per-lane parameters of one chunk (one receptor type) for one ligand atom.
struct ChunkPar {
  float p1a, p2a, pmf0, pmf1, pb, pc, ha, hb, hc;
  float a0, a1;  // p1a lj0, p1a lj0 lj1: den = a0 + a1 d^2 + d^9
  float nr2;  // contacts and the near electrostatic form lie within sqrt(nr2)
};

// Receptor points [p, p + NP) of one chunk against the lane's APL atoms.
// All terms of src/kernel_cuda_l2_calcenergy.cu:105-184 plus the contact
// count of CalcMcc_d:
//   vdW  (p1 - p2) / (p1 lj0 (1 + lj1 d^2) + 1), p1 = p1a/d^9, p2 = p2a/d^6,
//        evaluated multiplied through by d^9:
//        num/den = (p1a - p2a d^3) / (a0 + a1 d^2 + d^6 d^3);
//   ele  ele_p * g(el1 d) (the ligand charge is applied per atom later);
//        g = 1/(el1 d) is accumulated for every pair (the point stores
//        ele_p/el1); where el1 d < 1 the near form replaces it;
//   pmf  pmf1 / (1 + exp((6 - d/2)(d - pmf0))), exponent in base 2 as
//        A d^2 + B d + C; vdW and PMF share one reciprocal 1/(den (1 + e));
//        `pmf` here sums 1/(1 + e) over the chunk (pmf1 is applied per chunk);
//   hdb  -hdb1/sqrt(2 pi) exp(-((d - hdb0) hdb1)^2 / 2) (zero record if hdb0 <= 0.1);
//   hpc  hpp * switch(d^2) for d^2 <= 81 (KIND >= 1; S.hpp holds hpp / 2);
//   psp  psp[lig_t][seq3r] when class 2 and d <= pmf0 (KIND == 2);
//   contacts  d <= pmf0.
// Every pair takes the full form of every term, so a pair's contribution is
// a function of that pair alone. One wave-uniform branch skips work no lane
// needs: contacts, the near electrostatic form and PSP run only when some lane
// is within contact range (the terms inside test each pair exactly, so pairs
// outside add exact zeros). Two chunk-uniform choices are template parameters:
// CLAMP keeps the PMF exponent finite where the chunk's pmf0 lets it grow
// past 60, HB evaluates the H-bond term where some lane's atom has an H-bond
// record with the chunk's receptor type.
template <int APL, int KIND, int NP, bool CLAMP, bool HB>
__device__ __forceinline__ void pair_group(const Lds &S, int p, const float4 (&P)[NP],
                                           const PairConst &K, float nr2max,
                                           const float (&x)[APL], const float (&y)[APL],
                                           const float (&z)[APL], const ChunkPar (&cp)[APL],
                                           const float *const (&psprow)[APL],
                                           float (&vdw)[APL], float (&ele)[APL],
                                           float (&pmf)[APL], float (&psp)[APL],
                                           float (&hdb)[APL], float (&hpc)[APL],
                                           int (&cnt)[APL]) {
  const float pa = -0.5f * LOG2E;
  float hp[NP];
  int ps[NP];
#pragma unroll
  for (int j = 0; j < NP; ++j) {
    hp[j] = (KIND >= KIND_HPC) ? S.hpp[p + j] : 0.0f;
    ps[j] = (KIND == KIND_PSP) ? S.psp[p + j] : 0;
  }
  float d2[NP][APL], rd[NP][APL], dd[NP][APL], num[NP][APL], den[NP][APL];
  bool near = false;
#pragma unroll
  for (int j = 0; j < NP; ++j) {
    // contact range: the lane's nearest atom against the largest of its
    // atoms' near radii (a superset of the pairs the branch needs; the terms
    // inside test each pair exactly)
    float qmin = 0.0f;
#pragma unroll
    for (int k = 0; k < APL; ++k) {
      const float q2 = pair_d2(x[k], y[k], z[k], P[j].x, P[j].y, P[j].z);
      const float r = fast_rsq(q2);
      const float d = q2 * r;
      d2[j][k] = q2;
      rd[j][k] = r;
      dd[j][k] = d;
      qmin = k == 0 ? q2 : fminf(qmin, q2);
      const float d3 = q2 * d;
      const float d6 = d3 * d3;
      num[j][k] = fmaf(-cp[k].p2a, d3, cp[k].p1a);
      den[j][k] = fmaf(d6, d3, fmaf(cp[k].a1, q2, cp[k].a0));
      ele[k] = fmaf(P[j].w, r, ele[k]);
    }
    near |= qmin <= nr2max;
  }
#pragma unroll
  for (int j = 0; j < NP; ++j)
#pragma unroll
    for (int k = 0; k < APL; ++k) {
      const float q2 = d2[j][k], d = dd[j][k];
      // the exponent is clamped at 64 (exp2 must stay finite) only in chunks
      // whose parameters let it reach 60; elsewhere the clamp is a no-op
      const float wv0 = fmaf(pa, q2, fmaf(cp[k].pb, d, cp[k].pc));
      const float wv = CLAMP ? fminf(wv0, 64.0f) : wv0;
      // den (1 + e) and num (1 + e) as den + den e and num + num e: one
      // fma each, one rounding each
      const float e = fast_exp2(wv);
      const float r = fast_rcp(fmaf(den[j][k], e, den[j][k]));
      vdw[k] = fmaf(fmaf(num[j][k], e, num[j][k]), r, vdw[k]);
      pmf[k] = fmaf(den[j][k], r, pmf[k]);
      // H-bond: only in chunks where some lane's atom has an H-bond record
      // with this receptor type (elsewhere every term would be 0 * exp2)
      if (HB) {
        const float h = fmaf(d, cp[k].ha, cp[k].hb);
        hdb[k] = fmaf(cp[k].hc, fast_exp2(-(h * h)), hdb[k]);
      }
      if (KIND >= KIND_HPC) {
        // u <= 1, so clamping to [0, 1] is max(u, 0): the fma's own clamp
        const float u = __builtin_amdgcn_fmed3f(fmaf(q2, -INV81, 1.0f), 0.0f, 1.0f);
        const float u3 = u * u * u;
        hpc[k] = fmaf(hp[j], fmaf(u3, u, u3), hpc[k]);
      }
    }
  if (__ballot(near) != 0) {
#pragma unroll
    for (int j = 0; j < NP; ++j)
#pragma unroll
      for (int k = 0; k < APL; ++k) {
        const float d = dd[j][k];
        const bool c = d <= cp[k].pmf0;
        cnt[k] += c ? 1 : 0;
        const float s = K.el1 * d;
        if (s < 1.0f) {
          const float gn = fmaf(s * s, fmaf(K.b1, s, K.a1), K.el0);
          ele[k] = fmaf(P[j].w, fmaf(K.el1, gn, -rd[j][k]), ele[k]);
        }
        if (KIND == KIND_PSP) {
          if (c) psp[k] += psprow[k][ps[j]];
        }
      }
  }
}

template <int APL, int KIND, bool CLAMP, bool HB>
__device__ __forceinline__ void chunk_pairs(const Lds &S, int beg, int end, const PairConst &K,
                                            float nr2max,
                                            const float (&x)[APL], const float (&y)[APL],
                                            const float (&z)[APL], const ChunkPar (&cp)[APL],
                                            const float *const (&psprow)[APL],
                                            float (&vdw)[APL], float (&ele)[APL],
                                            float (&pmf)[APL], float (&psp)[APL],
                                            float (&hdb)[APL], float (&hpc)[APL],
                                            int (&cnt)[APL]) {
  // points per group: enough independent pairs per lane to cover latency
  constexpr int NP = APL >= 3 ? 1 : (APL == 2 ? 2 : 4);
  int p = beg;
  // the chunk's PMF sum is taken without pmf1, which multiplies it once at
  // the end
  float pmfc[APL];
#pragma unroll
  for (int k = 0; k < APL; ++k) pmfc[k] = 0.0f;
  // the next group's points are read from LDS one group ahead, so the read
  // latency overlaps the current group's arithmetic (indices clamped into
  // the chunk: the last prefetch re-reads its final point)
  // two register sets in turn (no copies): while one group is evaluated the
  // other is being read
  if (p + NP <= end) {
    float4 Pa[NP], Pb[NP];
#pragma unroll
    for (int j = 0; j < NP; ++j) Pa[j] = S.xyze[p + j];
    for (; p + 2 * NP <= end; p += 2 * NP) {
#pragma unroll
      for (int j = 0; j < NP; ++j) Pb[j] = S.xyze[p + NP + j];
      pair_group<APL, KIND, NP, CLAMP, HB>(S, p, Pa, K, nr2max, x, y, z, cp, psprow, vdw, ele,
                                           pmfc, psp, hdb, hpc, cnt);
#pragma unroll
      for (int j = 0; j < NP; ++j) Pa[j] = S.xyze[min(p + 2 * NP + j, end - 1)];
      pair_group<APL, KIND, NP, CLAMP, HB>(S, p + NP, Pb, K, nr2max, x, y, z, cp, psprow, vdw,
                                           ele, pmfc, psp, hdb, hpc, cnt);
    }
    if (p + NP <= end) {
      pair_group<APL, KIND, NP, CLAMP, HB>(S, p, Pa, K, nr2max, x, y, z, cp, psprow, vdw, ele,
                                           pmfc, psp, hdb, hpc, cnt);
      p += NP;
    }
  }
  for (; p < end; ++p) {
    const float4 P1[1] = {S.xyze[p]};
    pair_group<APL, KIND, 1, CLAMP, HB>(S, p, P1, K, nr2max, x, y, z, cp, psprow, vdw, ele, pmfc,
                                        psp, hdb, hpc, cnt);
  }
#pragma unroll
  for (int k = 0; k < APL; ++k) pmf[k] = fmaf(cp[k].pmf1, pmfc[k], pmf[k]);
}

// Sum over MCS templates [m_begin, m_end) of tcc_m * sqrt(mean squared
// distance of the template's atoms), src/kernel_cuda_l2_calcenergy.cu:274-313. Templates go
// four at a time: their loads are issued together and their lane sums share
// each butterfly step, so the per-template latencies overlap. Per template
// the additions are the same as one template at a time.
template <int APL>
__device__ __forceinline__ float mcs_sum(const DevComplex &C, int lpr, const float (&x)[APL],
                                         const float (&y)[APL], const float (&z)[APL],
                                         const int (&li)[APL], const bool (&val)[APL],
                                         int m_begin, int m_end) {
  float elhm = 0.0f;
  const int pos = m_end;
  for (int m0 = m_begin; m0 < pos; m0 += 4) {
    float s[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      s[q] = 0.0f;
      const int m = m0 + q;
      if (m >= pos) continue;
#pragma unroll
      for (int k = 0; k < APL; ++k) {
        if (!val[k]) continue;
        const float4 M = C.mcs[(size_t)m * C.lna + li[k]];
        if (M.w != 0.0f) s[q] += pair_d2(x[k], y[k], z[k], M.x, M.y, M.z);
      }
    }
    for (int b = 1; b < lpr; b <<= 1) {
#pragma unroll
      for (int q = 0; q < 4; ++q) s[q] += __shfl_xor(s[q], b);
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int m = m0 + q;
      if (m >= pos) break;
      const float n = C.mcs_cnt[m];
      if (n > 0.0f) elhm += C.mcs_tcc[m] * sqrtf(s[q] / n);
    }
  }
  return elhm;
}

// Energy, RMSD and MCC of one pose held by `lpr` lanes (CalcEnergy_d +
// CombineEnergy_d + CalcRmsd_d + CalcMcc_d). Every lane of the replica ends
// with the same bit pattern in E.
//
// With a receptor split (sp.ns > 1) every thread of the workgroup must call
// this together: `live` (wave-uniform) says whether the wave has a pose to
// evaluate; the groups exchange their partial sums through LDS behind one
// workgroup barrier and add them in group order, so every group ends with
// identical bits (with two groups this is the order-free g0 + g1).
//
// TILE: the receptor does not stay resident in LDS; its sorted points stream
// through two LDS tile buffers of S.tile points. While a tile is evaluated
// the next one is read from HBM into registers (one point per thread; points
// beyond the workgroup size after the evaluation) and written to the other
// buffer; one workgroup barrier per tile.
// Every thread runs the tile loop (waves without a pose only stage). Chunks
// that cross a tile boundary are evaluated as one fragment per tile.
template <int APL, bool TILE>
__device__ __forceinline__ void pose_energy(const DevComplex &C, const Lds &S, int lpr, Split &sp,
                            bool live, const float (&x)[APL],
                            const float (&y)[APL], const float (&z)[APL],
                            const float (&ox)[APL], const float (&oy)[APL],
                            const float (&oz)[APL], const int (&lt)[APL],
                            const float (&q)[APL], const int (&li)[APL],
                            const bool (&val)[APL], const float mm[6], gd_energy &E) {
  const int n_chunks = __builtin_amdgcn_readfirstlane(C.n_chunks);
  PairConst K;
  K.el0 = C.el0;
  K.el1 = C.el1;
  K.a1 = C.a1;
  K.b1 = C.b1;

  float vdw[APL], ele[APL], pmf[APL], psp[APL], hdb[APL], hpc[APL];
  int cnt[APL];
  const float *psprow[APL];
#pragma unroll
  for (int k = 0; k < APL; ++k) {
    vdw[k] = ele[k] = pmf[k] = psp[k] = hdb[k] = hpc[k] = 0.0f;
    cnt[k] = 0;
    psprow[k] = C.psp + (size_t)lt[k] * C.n_res;
  }

  // fragment [fb, fe) of chunk ci, indices into T's point arrays; each lane
  // group of a receptor split takes its share
  auto fragment = [&](const Lds &T, int ci, int fb, int fe) {
    const int4 ch = S.ch[ci];
    const int fl = fe - fb;
    const int beg = fb + (fl * sp.h) / sp.ns;
    const int end = fb + (fl * (sp.h + 1)) / sp.ns;
    const int type = __builtin_amdgcn_readfirstlane(ch.z);
    const int kind = __builtin_amdgcn_readfirstlane(ch.w);
    ChunkPar cp[APL];
#pragma unroll
    for (int k = 0; k < APL; ++k) {
      const float *r = S.par + (type * NLT + lt[k]) * NPAR;
      // the record holds the chunk constants precomputed on the host:
      // a0 = p1a lj0, a1 = p1a lj0 lj1; (6 - d/2)(d - pmf0) log2(e) =
      // A d^2 + B d + C with B = log2(e)(6 + pmf0/2), C = -6 log2(e) pmf0;
      // contacts (d <= pmf0) and the near electrostatic form (d < 1/el1) lie
      // within sqrt(nr2), nr2 = (1.0001 max(pmf0, 1/el1))^2 (the margin covers
      // the rounding of d)
      const float4 a = *reinterpret_cast<const float4 *>(r);
      const float4 b = *reinterpret_cast<const float4 *>(r + 4);
      const float4 c = *reinterpret_cast<const float4 *>(r + 8);
      cp[k].p1a = a.x;
      cp[k].p2a = a.y;
      cp[k].pmf0 = a.z;
      cp[k].pmf1 = a.w;
      cp[k].ha = b.x;
      cp[k].hb = b.y;
      cp[k].hc = b.z;
      cp[k].nr2 = b.w;
      cp[k].a0 = c.x;
      cp[k].a1 = c.y;
      cp[k].pb = c.z;
      cp[k].pc = c.w;
    }
    // the PMF exponent (6 - d/2)(d - pmf0) log2(e) peaks at (pmf0 - 12)^2 / 8
    // log2(e); only a chunk where some lane's peak passes 60 needs the clamp
    bool wide = false, hbond = false;
    float nr2max = cp[0].nr2;
#pragma unroll
    for (int k = 0; k < APL; ++k) {
      const float t = cp[k].pmf0 - 12.0f;
      wide |= !(0.125f * LOG2E * t * t <= 60.0f);
      hbond |= cp[k].hc != 0.0f;
      nr2max = fmaxf(nr2max, cp[k].nr2);
    }
    const bool clamp = __ballot(wide) != 0;
    const bool hb = __ballot(hbond) != 0;
    auto pairs = [&](auto kind_c, auto clamp_c, auto hb_c) {
      chunk_pairs<APL, decltype(kind_c)::value, decltype(clamp_c)::value,
                  decltype(hb_c)::value>(T, beg, end, K, nr2max, x, y, z, cp, psprow, vdw, ele,
                                         pmf, psp, hdb, hpc, cnt);
    };
    auto by_hb = [&](auto kind_c, auto clamp_c) {
      if (hb) pairs(kind_c, clamp_c, std::true_type{});
      else pairs(kind_c, clamp_c, std::false_type{});
    };
    auto by_clamp = [&](auto kind_c) {
      if (clamp) by_hb(kind_c, std::true_type{});
      else by_hb(kind_c, std::false_type{});
    };
    if (kind == KIND_PLAIN) by_clamp(std::integral_constant<int, KIND_PLAIN>{});
    else if (kind == KIND_HPC) by_clamp(std::integral_constant<int, KIND_HPC>{});
    else by_clamp(std::integral_constant<int, KIND_PSP>{});
  };

  if constexpr (!TILE) {
    for (int ci = 0; live && ci < n_chunks; ++ci) {
      const int4 ch = S.ch[ci];
      const int cb = __builtin_amdgcn_readfirstlane(ch.x);
      const int cl = __builtin_amdgcn_readfirstlane(ch.y);
      fragment(S, ci, cb, cb + cl);
    }
  } else {
    const int T = S.tile, pnp = C.pnp;
    const int tid = threadIdx.x, nt = blockDim.x;
    const int n_tiles = (pnp + T - 1) / T;
    for (int i = tid; i < min(T, pnp); i += nt) {
      S.tx[i] = S.gx[i];
      S.th[i] = 0.5f * S.ghpp[i];
      S.tp[i] = S.gpsp[i];
    }
    __syncthreads();
    int ci = 0;
    for (int t = 0; t < n_tiles; ++t) {
      const int tb = t * T, te = min(tb + T, pnp);
      // this thread's point of the next tile, read while this tile is evaluated
      const int ni = te + tid;
      const bool stage = tid < T && ni < min(te + T, pnp);
      float4 nx = make_float4(0.f, 0.f, 0.f, 0.f);
      float nh = 0.0f;
      int np = 0;
      if (stage) {
        nx = S.gx[ni];
        nh = S.ghpp[ni];
        np = S.gpsp[ni];
      }
      if (live) {
        Lds Tl = S;
        const int cur = (t & 1) * T;
        Tl.xyze = S.tx + cur;
        Tl.hpp = S.th + cur;
        Tl.psp = S.tp + cur;
        while (ci < n_chunks) {
          const int4 ch = S.ch[ci];
          const int cb = __builtin_amdgcn_readfirstlane(ch.x);
          const int ce = cb + __builtin_amdgcn_readfirstlane(ch.y);
          const int fb = max(cb, tb), fe = min(ce, te);
          if (fb < fe) fragment(Tl, ci, fb - tb, fe - tb);
          if (ce > te) break;  // the chunk continues in the next tile
          ++ci;
        }
      }
      if (stage) {
        const int o = ((t + 1) & 1) * T + tid;
        S.tx[o] = nx;
        S.th[o] = 0.5f * nh;
        S.tp[o] = np;
      }
      // a tile longer than the workgroup: the rest of it, read now
      for (int i = ni + nt; i < min(te + T, pnp); i += nt) {
        const int o = ((t + 1) & 1) * T + (i - te);
        S.tx[o] = S.gx[i];
        S.th[o] = 0.5f * S.ghpp[i];
        S.tp[o] = S.gpsp[i];
      }
      __syncthreads();
    }
  }

  // Per-atom lists that the groups of a receptor split share out the way they
  // share the receptor's chunks: each group takes its part of the KDE points
  // of every atom's type, of every atom's reference contacts and of the MCS
  // templates, and the partial sums join the exchange below. With one group
  // the parts are the whole lists.
  const float kde2l = C.kde2l;
  float t_kde = 0.f, elhm = 0.f;
  int t_tp = 0;
  if (live) {
#pragma unroll
    for (int k = 0; k < APL; ++k) {
      if (!val[k]) continue;
      // KDE: mean over same-type points, src/kernel_cuda_l2_calcenergy.cu:224-253
      const int2 kr = S.kr[lt[k]];
      const int kl = kr.y - kr.x;
      const int kb = kr.x + (kl * sp.h) / sp.ns, ke = kr.x + (kl * (sp.h + 1)) / sp.ns;
      const float ks = kde_sum(S.kde, kde2l, x[k], y[k], z[k], kb, ke);
      if (kl > 0) t_kde += ks / (float)kl;
      // true positives against the reference contact list
      const int nr = C.ref_cnt[li[k]];
      const int2 *rl = C.ref + (size_t)li[k] * KMAX;
      const int rb0 = (nr * sp.h) / sp.ns, re0 = (nr * (sp.h + 1)) / sp.ns;
      int j = rb0;
      // four contacts at a time: their list entries, then their points, are
      // read together (the count is order-free)
      for (; j + 4 <= re0; j += 4) {
        int2 r[4];
        float4 P[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) r[u] = rl[j + u];
#pragma unroll
        for (int u = 0; u < 4; ++u) P[u] = TILE ? S.gx[r[u].x] : S.xyze[r[u].x];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const float d2 = pair_d2(x[k], y[k], z[k], P[u].x, P[u].y, P[u].z);
          const float d = d2 * fast_rsq(d2);
          t_tp += (d <= __int_as_float(r[u].y)) ? 1 : 0;
        }
      }
      for (; j < re0; ++j) {
        const int2 r = rl[j];
        const float4 P = TILE ? S.gx[r.x] : S.xyze[r.x];
        const float d2 = pair_d2(x[k], y[k], z[k], P.x, P.y, P.z);
        const float d = d2 * fast_rsq(d2);
        t_tp += (d <= __int_as_float(r.y)) ? 1 : 0;
      }
    }
    // MCS restraints, src/kernel_cuda_l2_calcenergy.cu:274-313
    elhm = mcs_sum<APL>(C, lpr, x, y, z, li, val, (C.pos * sp.h) / sp.ns,
                        (C.pos * (sp.h + 1)) / sp.ns);
  }

  // pair sums that are linear in the receptor points
  float t_vdw = 0.f, t_ele = 0.f, t_pmf = 0.f, t_psp = 0.f, t_hdb = 0.f, t_hpc = 0.f;
  float t_rms = 0.f;
  int t_cnt = 0;
#pragma unroll
  for (int k = 0; k < APL; ++k) {
    if (!val[k]) continue;
    t_vdw += vdw[k];
    t_ele += q[k] * ele[k];
    t_pmf += pmf[k];
    t_psp += psp[k];
    t_hdb += hdb[k];
    t_cnt += cnt[k];
  }
  t_vdw = seg_sum(t_vdw, lpr);
  t_ele = seg_sum(t_ele, lpr);
  t_pmf = seg_sum(t_pmf, lpr);
  t_psp = seg_sum(t_psp, lpr);
  t_hdb = seg_sum(t_hdb, lpr);
  t_cnt = seg_sum_i(t_cnt, lpr);
  t_kde = seg_sum(t_kde, lpr);
  t_tp = seg_sum_i(t_tp, lpr);
  if (sp.ns > 1) {
    // exchange: per-atom HPC sums (the restraint is nonlinear in them) and
    // the replica's linear sums. Every group adds the ns partials in group
    // order, so all groups end with the same bits.
    float *buf = S.xp + sp.par * sp.xp_floats;
    float *hb = buf;                                // [ns*half][4]
    float *rb = buf + (size_t)sp.ns * sp.half * 4;  // [ns][slots][12]
    if (live) {
      float hg[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int k = 0; k < APL; ++k) hg[k] = hpc[k];
      *reinterpret_cast<float4 *>(hb + (size_t)threadIdx.x * 4) = make_float4(hg[0], hg[1], hg[2], hg[3]);
      if ((sp.ltid - sp.slot * lpr) == 0) {
        float4 *r = reinterpret_cast<float4 *>(rb + ((size_t)sp.h * sp.slots + sp.slot) * 12);
        r[0] = make_float4(t_vdw, t_ele, t_pmf, t_psp);
        r[1] = make_float4(t_hdb, __int_as_float(t_cnt), t_kde, __int_as_float(t_tp));
        r[2] = make_float4(elhm, 0.f, 0.f, 0.f);
      }
    }
    __syncthreads();
    if (live) {
      for (int g = 0; g < sp.ns; ++g) {
        const float4 hv = *reinterpret_cast<const float4 *>(hb + ((size_t)g * sp.half + sp.ltid) * 4);
        const float4 *r = reinterpret_cast<const float4 *>(rb + ((size_t)g * sp.slots + sp.slot) * 12);
        const float4 r0 = r[0], r1 = r[1], r2 = r[2];
        const float hg[4] = {hv.x, hv.y, hv.z, hv.w};
        if (g == 0) {
#pragma unroll
          for (int k = 0; k < APL; ++k) hpc[k] = hg[k];
          t_vdw = r0.x;
          t_ele = r0.y;
          t_pmf = r0.z;
          t_psp = r0.w;
          t_hdb = r1.x;
          t_cnt = __float_as_int(r1.y);
          t_kde = r1.z;
          t_tp = __float_as_int(r1.w);
          elhm = r2.x;
        } else {
#pragma unroll
          for (int k = 0; k < APL; ++k) hpc[k] += hg[k];
          t_vdw += r0.x;
          t_ele += r0.y;
          t_pmf += r0.z;
          t_psp += r0.w;
          t_hdb += r1.x;
          t_cnt += __float_as_int(r1.y);
          t_kde += r1.z;
          t_tp += __float_as_int(r1.w);
          elhm += r2.x;
        }
      }
    }
    sp.par ^= 1;
  }
  if (!live) return;

  // per-atom terms
#pragma unroll
  for (int k = 0; k < APL; ++k) {
    if (!val[k]) continue;
    // hydrophobic restraint, src/kernel_cuda_l2_calcenergy.cu:200-205
    const float h2 = (hpc[k] - C.hpl0[lt[k]]) / C.hpl1[lt[k]];
    t_hpc += 0.5f * h2 * h2 - C.hpl2[lt[k]];
    // RMSD against the start pose
    const float dx = x[k] - (ox[k] + C.pocket[0]);
    const float dy = y[k] - (oy[k] + C.pocket[1]);
    const float dz = z[k] - (oz[k] + C.pocket[2]);
    t_rms += dx * dx + dy * dy + dz * dz;
  }
  const int pos = C.pos;

  t_hpc = seg_sum(t_hpc, lpr);
  t_rms = seg_sum(t_rms, lpr);

  // src/kernel_cuda_l2_calcenergy.cu:334-359, src/kernel_cuda_l3_combineenergy.cu:28-37
  const float lna = (float)C.lna;
  float e[GD_MAXWEI];
  e[0] = t_vdw / lna;
  e[1] = t_ele / lna;
  e[2] = t_pmf / lna;
  e[3] = t_psp / lna;
  e[4] = t_hdb / lna;
  e[5] = t_hpc / lna;
  e[6] = (t_kde / C.kde3) / lna;
  e[7] = pos > 0 ? logf(elhm / (float)pos) : 0.0f;
  {
    const float cx = C.pocket[0] + mm[0], cy = C.pocket[1] + mm[1], cz = C.pocket[2] + mm[2];
    const float dx = cx - C.pocket[0], dy = cy - C.pocket[1], dz = cz - C.pocket[2];
    e[8] = sqrtf(dx * dx + dy * dy + dz * dz);
  }
  float total = 0.0f;
#pragma unroll
  for (int i = 0; i < GD_MAXWEI - 1; ++i) {
    e[i] = C.a_para[i] * e[i] + C.b_para[i];
    E.e[i] = e[i];
    total += e[i] * C.w[i];
  }
  E.e[GD_MAXWEI - 1] = total;
  E.rmsd = sqrtf(t_rms / lna);
  // Matthews correlation of the contact maps, src/kernel_cuda_l2_calcmcc.cu:194-206
  const int tp = t_tp, fp = t_cnt - t_tp, fn = C.ref_total - t_tp;
  const int tn = C.lna * C.pnp - C.ref_total - fp;
  const float tp0 = (float)tp, fn0 = (float)fn, fp0 = (float)fp, tn0 = (float)tn;
  const float dv = sqrtf((tp0 + fp0) * (tp0 + fn0) * (tn0 + fp0) * (tn0 + fn0));
  E.cms = dv != 0.0f ? (tp0 * tn0 - fp0 * fn0) / dv : 0.0f;
}

// Weighted KDE, MCS and pocket-distance terms (e[6..8] after normalization)
// of one pose: the cheap part of the energy, evaluated exactly inside the
// surrogate (src/kernel_cuda_l2_calcenergy.cu:224-326).
//
// SHARE (a split pack of a complex with large KDE or MCS sets): the groups of
// the receptor split take their parts of each atom's KDE points and of the
// MCS templates and exchange the partial sums through the split's LDS
// buffers behind one barrier, adding them in group order. Every thread of the
// workgroup must then call this together; `live` (wave-uniform) says whether
// the wave has a pose. Without SHARE the lists are evaluated whole by every
// group, and only waves with a pose call it.
template <int APL, bool SHARE>
__device__ __forceinline__ float cheap_terms(const DevComplex &C, const Lds &S, int lpr,
                                             Split &sp, bool live,
                                             const float (&x)[APL], const float (&y)[APL],
                                             const float (&z)[APL], const int (&lt)[APL],
                                             const bool (&val)[APL], const float mm[6],
                                             const Rot &R, int a, int ilig) {
  const int gh = SHARE ? sp.h : 0, gn = SHARE ? sp.ns : 1;
  const int pos = C.pos;
  float t_kde = 0.0f, elhm = 0.0f;
  if (live) {
    const float kde2l = C.kde2l;
#pragma unroll
    for (int k = 0; k < APL; ++k) {
      if (!val[k]) continue;
      const int2 kr = S.kr[lt[k]];
      const int kl = kr.y - kr.x;
      const float ks = kde_sum(S.kde, kde2l, x[k], y[k], z[k], kr.x + (kl * gh) / gn,
                               kr.x + (kl * (gh + 1)) / gn);
      if (kl > 0) t_kde += ks / (float)kl;
    }
    // MCS restraints in closed form from the conformer's template moments:
    // sum |R o + t - m|^2 = A + n |R obar + t - mbar|^2 - 2 <R, M>; the
    // replica's lanes take the group's templates in turn
    const float4 *qm = C.mcsq + (size_t)ilig * pos * 4;
    const float tx = mm[0] + C.pocket[0], ty = mm[1] + C.pocket[1], tz = mm[2] + C.pocket[2];
    const int mb = (pos * gh) / gn, me = (pos * (gh + 1)) / gn;
    for (int m = mb + a; m < me; m += lpr) {
      const float n = C.mcs_cnt[m];
      if (!(n > 0.0f)) continue;
      const float4 q0 = qm[4 * m], q1 = qm[4 * m + 1], q2 = qm[4 * m + 2], q3 = qm[4 * m + 3];
      const float ux = R.r00 * q0.y + R.r01 * q0.z + R.r02 * q0.w + tx - q1.x;
      const float uy = R.r10 * q0.y + R.r11 * q0.z + R.r12 * q0.w + ty - q1.y;
      const float uz = R.r20 * q0.y + R.r21 * q0.z + R.r22 * q0.w + tz - q1.z;
      const float rm = R.r00 * q1.w + R.r01 * q2.x + R.r02 * q2.y + R.r10 * q2.z +
                       R.r11 * q2.w + R.r12 * q3.x + R.r20 * q3.y + R.r21 * q3.z + R.r22 * q3.w;
      const float sq = fmaxf(q0.x + n * (ux * ux + uy * uy + uz * uz) - 2.0f * rm, 0.0f);
      elhm += C.mcs_tcc[m] * sqrtf(sq / n);
    }
    elhm = seg_sum(elhm, lpr);
    t_kde = seg_sum(t_kde, lpr);
  }
  if constexpr (SHARE) {
    float *rb = S.xp + sp.par * sp.xp_floats + (size_t)sp.ns * sp.half * 4;  // [ns][slots][12]
    if (live && (sp.ltid - sp.slot * lpr) == 0)
      *reinterpret_cast<float4 *>(rb + ((size_t)sp.h * sp.slots + sp.slot) * 12) =
          make_float4(t_kde, elhm, 0.f, 0.f);
    __syncthreads();
    if (live) {
      for (int g = 0; g < sp.ns; ++g) {
        const float4 r = *reinterpret_cast<const float4 *>(rb + ((size_t)g * sp.slots + sp.slot) * 12);
        t_kde = g == 0 ? r.x : t_kde + r.x;
        elhm = g == 0 ? r.y : elhm + r.y;
      }
    }
    sp.par ^= 1;
  }
  if (!live) return 0.0f;
  const float lna = (float)C.lna;
  const float e6 = (t_kde / C.kde3) / lna;
  const float e7 = pos > 0 ? logf(elhm / (float)pos) : 0.0f;
  const float cx = C.pocket[0] + mm[0], cy = C.pocket[1] + mm[1], cz = C.pocket[2] + mm[2];
  const float dx = cx - C.pocket[0], dy = cy - C.pocket[1], dz = cz - C.pocket[2];
  const float e8 = sqrtf(dx * dx + dy * dy + dz * dz);
  return (C.a_para[6] * e6 + C.b_para[6]) * C.w[6] + (C.a_para[7] * e7 + C.b_para[7]) * C.w[7] +
         (C.a_para[8] * e8 + C.b_para[8]) * C.w[8];
}

// Surrogate total of the pose held by each replica of the wave: the MLP's
// receptor-pair estimate, scale * sum_atoms y / lna, plus the exact cheap
// terms. Rows are the
// wave's atoms, 16 per MFMA tile: tile (k, sub) holds atom k of lanes
// 16 sub .. 16 sub + 15. `da` is the wave's scratch: X tile [16][64] bf16,
// A1 tile [16][128] bf16, y[16].
template <int APL, int DAM, bool SHARE>
__device__ __forceinline__ float sur_total(const KParams &P, const DevComplex &C, const Lds &S,
                                           Split &sp, bool live,
                                           unsigned char *da, int lpr, int prt,
                                           const float (&x)[APL], const float (&y)[APL],
                                           const float (&z)[APL], const int (&lt)[APL],
                                           const bool (&val)[APL], const float mm[6],
                                           const Rot &R, int a, int ilig) {
  if (!live) {  // the cheap terms' exchange only (a workgroup barrier)
    if constexpr (SHARE)
      (void)cheap_terms<APL, true>(C, S, lpr, sp, false, x, y, z, lt, val, mm, R, a, ilig);
    return 0.0f;
  }
  if constexpr (DAM == 2) {
    // tabulated network: trilinear interpolation of the atom type's output
    // grid (same grid geometry and clamping as the feature grid)
    const int n = C.sn;
    const size_t n3 = (size_t)n * n * n;
    const float lim = (float)(n - 1) - 1e-4f;
    const float ih = 1.0f / C.sh;
    float ys = 0.0f;
#pragma unroll
    for (int k = 0; k < APL; ++k) {
      const float ux = fminf(fmaxf((x[k] - C.sox) * ih, 0.0f), lim);
      const float uy = fminf(fmaxf((y[k] - C.soy) * ih, 0.0f), lim);
      const float uz = fminf(fmaxf((z[k] - C.soz) * ih, 0.0f), lim);
      const int ix = (int)ux, iy = (int)uy, iz = (int)uz;
      const float tx = ux - ix, ty = uy - iy, tz = uz - iz;
      const float *g = C.stab + ((size_t)prt * NLT + lt[k]) * n3 + ((size_t)ix * n + iy) * n + iz;
      const size_t sx = (size_t)n * n, sy = n;
      const float c00 = fmaf(tz, g[1] - g[0], g[0]);
      const float c01 = fmaf(tz, g[sy + 1] - g[sy], g[sy]);
      const float c10 = fmaf(tz, g[sx + 1] - g[sx], g[sx]);
      const float c11 = fmaf(tz, g[sx + sy + 1] - g[sx + sy], g[sx + sy]);
      const float c0 = fmaf(ty, c01 - c00, c00);
      const float c1 = fmaf(ty, c11 - c10, c10);
      const float v = fmaf(tx, c1 - c0, c0);
      ys += val[k] ? v : 0.0f;
    }
    ys = seg_sum(ys, lpr);
    return P.sur_scale * (ys / (float)C.lna) +
           cheap_terms<APL, SHARE>(C, S, lpr, sp, true, x, y, z, lt, val, mm, R, a, ilig);
  }
  const int l = threadIdx.x & 63;
  unsigned short *Xs = reinterpret_cast<unsigned short *>(da);
  unsigned short *A1s = Xs + 16 * SUR_F;
  float *ys = reinterpret_cast<float *>(A1s + 16 * SUR_H);
  SurW W;
  W.W1T = P.sW1T;
  W.W2T = P.sW2T;
  W.W2R = nullptr;
  W.b2 = P.sb2;
  W.w3 = P.sw3;
  W.b3 = P.sb3;
  SurGrid G;
  G.grid = C.sgrid;
  G.ox = C.sox;
  G.oy = C.soy;
  G.oz = C.soz;
  G.h = C.sh;
  G.n = C.sn;
  float ysum = 0.0f;
#pragma unroll
  for (int k = 0; k < APL; ++k) {
    float f[32];
    grid_radial(G, prt, x[k], y[k], z[k], f);
    for (int sub = 0; sub < 4; ++sub) {
      if ((l >> 4) == sub) {
        unsigned short *row = Xs + (l & 15) * SUR_F;
#pragma unroll
        for (int q = 0; q < NLT; ++q) row[q] = (q == lt[k]) ? (unsigned short)0x3F80 : (unsigned short)0;
#pragma unroll
        for (int q = 0; q < 32; ++q) row[NLT + q] = f2bf(f[q]);
        row[SUR_BIAS_COL] = 0x3F80;  // 1.0
#pragma unroll
        for (int q = SUR_BIAS_COL + 1; q < SUR_F; ++q) row[q] = 0;
      }
      __builtin_amdgcn_s_waitcnt(0xc07f);
      __builtin_amdgcn_wave_barrier();
      f32x4_t h1[8], h2[8];
      mlp_fwd16(Xs, 0, W, A1s, h1, h2);
      float yy[4];
      mlp_out16(h2, W, yy);
      if ((l & 15) == 0)
#pragma unroll
        for (int i = 0; i < 4; ++i) ys[4 * (l >> 4) + i] = yy[i];
      __builtin_amdgcn_s_waitcnt(0xc07f);
      __builtin_amdgcn_wave_barrier();
      if ((l >> 4) == sub && val[k]) ysum += ys[l & 15];
      __builtin_amdgcn_s_waitcnt(0xc07f);
      __builtin_amdgcn_wave_barrier();
    }
  }
  ysum = seg_sum(ysum, lpr);
  // the network sums per-atom pair energies; the terms are per-atom means
  return P.sur_scale * (ysum / (float)C.lna) +
         cheap_terms<APL, SHARE>(C, S, lpr, sp, true, x, y, z, lt, val, mm, R, a, ilig);
}

// ExchangeReplicas_d for one ladder (src/kernel_cuda_l1_exchangereplicas.cu:33-78):
// temperature i1 = 2 pair + mode_t and i1 + 1 trade slots when
// exp((mb1 - mb2)(E2 - E1)) > U.
template <class U>
__device__ __forceinline__ void exchange_ladder(int n_tmp, int mode_t, const float *et, const float *mb,
                                int *temps, int *acs, int *order, U &&uni) {
  for (int t = 0; t < n_tmp; ++t) order[temps[t]] = t;
  const int maxpair = n_tmp / 2 - ((n_tmp % 2 == 0) && (mode_t == 1));
  for (int pair = 0; pair < maxpair; ++pair) {
    const int i1 = (pair << 1) + mode_t;
    const int i2 = i1 + 1;
    const int o1 = order[i1];
    const int o2 = order[i2];
    const int tt = temps[o1];
    const float e1 = et[o1], e2 = et[o2];
    const float mb1 = mb[temps[o1]], mb2 = mb[temps[o2]];
    const float d1 = mb1 - mb2;
    const float d2 = e2 - e1;
    const float delta = d1 * d2;
    const float prob = decision_expf(delta);
    if (prob > uni(pair)) {
      temps[o1] = temps[o2];
      temps[o2] = tt;
      acs[o1] += 1;
      acs[o2] += 1;
    }
  }
}

// DA: 0 plain Metropolis; delayed acceptance with the surrogate from the
// network evaluated per atom (1) or from its tabulated outputs (2); 3 and 4
// are 1 and 2 with the split groups' shared surrogate sums compiled in.
template <int APL, int DA, bool TILE>
__device__ __forceinline__ void run_pack(const KParams &P, const Pack &pk, const DevComplex &C, const Lds &S,
                         int lpr, int step0) {
  constexpr int DAM = DA > 2 ? DA - 2 : DA;
  constexpr bool SHARED = DA > 2;
  const int tid = threadIdx.x;
  Split sp;
  sp.ns = P.unsplit ? 1 : pk.ns;
  sp.half = P.unsplit ? (int)blockDim.x : pk.half;
  // half is a multiple of 64, so the group index is wave-uniform
  sp.h = __builtin_amdgcn_readfirstlane(tid / sp.half);
  sp.ltid = tid - sp.h * sp.half;
  sp.slots = sp.half / 8;
  sp.xp_floats = P.xp_floats;
  sp.par = 0;
  const int slot = sp.ltid / lpr;
  sp.slot = slot;
  const int a = sp.ltid - slot * lpr;
  const bool lead = a == 0 && sp.h == 0;  // writes the replica's results
  const int n_slots = pk.n_slots;
  const bool active = slot < n_slots && sp.h < sp.ns;
  const int sl = active ? slot : 0;
  const int sv = P.slots[pk.slot_off + sl];  // replica id or pose index
  const int lna = C.lna;
  const int n_lig = C.n_lig;

  int flat = 0, ilig;
  float mo[6];
  gd_energy eo;
  if (P.mode == MODE_EVAL || P.mode == MODE_SUREVAL) {
    ilig = P.ev_lig[sv];
#pragma unroll
    for (int i = 0; i < 6; ++i) mo[i] = P.ev_mm[(size_t)sv * 6 + i];
  } else {
    flat = P.rep_flat[sv];
    ilig = flat % n_lig;
#pragma unroll
    for (int i = 0; i < 6; ++i) mo[i] = P.mm[(size_t)sv * 6 + i];
    eo = P.e_old[sv];
  }

  // the lane's atoms: l = a + k*lpr
  float ox[APL], oy[APL], oz[APL], q[APL];
  int lt[APL], li[APL];
  bool val[APL];
  const float *orig = C.lig_orig + (size_t)ilig * 3 * lna;
#pragma unroll
  for (int k = 0; k < APL; ++k) {
    const int l = a + k * lpr;
    val[k] = l < lna;
    const int lc = val[k] ? l : 0;
    li[k] = lc;
    ox[k] = orig[lc];
    oy[k] = orig[lna + lc];
    oz[k] = orig[2 * lna + lc];
    lt[k] = C.lig_t[lc];
    q[k] = C.lig_c[lc];
  }

  float x[APL], y[APL], z[APL];
  Rot Rp{};  // rotation of the last placed pose (the surrogate's MCS moments use it)
  auto place = [&](const float m[6]) {
    const Rot R = euler(m[3], m[4], m[5]);
    Rp = R;
#pragma unroll
    for (int k = 0; k < APL; ++k) {
      x[k] = R.r00 * ox[k] + R.r01 * oy[k] + R.r02 * oz[k] + m[0] + C.pocket[0];
      y[k] = R.r10 * ox[k] + R.r11 * oy[k] + R.r12 * oz[k] + m[1] + C.pocket[1];
      z[k] = R.r20 * ox[k] + R.r21 * oy[k] + R.r22 * oz[k] + m[2] + C.pocket[2];
    }
  };

  if constexpr (DA) {
    if (P.mode == MODE_SUREVAL) {
      extern __shared__ __attribute__((aligned(16))) unsigned char lds_se[];
      unsigned char *dse = lds_se + P.off_da + (threadIdx.x >> 6) * 6208;
      place(mo);
      const bool live = __ballot(active) != 0;
      float sv_s;
      if constexpr (SHARED) {
        sv_s = (sp.ns > 1 && C.da_share != 0)
                   ? sur_total<APL, DAM, true>(P, C, S, sp, live, dse, lpr, pk.prt, x, y, z, lt,
                                               val, mo, Rp, a, ilig)
                   : sur_total<APL, DAM, false>(P, C, S, sp, live, dse, lpr, pk.prt, x, y, z, lt,
                                                val, mo, Rp, a, ilig);
      } else {
        sv_s = sur_total<APL, DAM, false>(P, C, S, sp, live, dse, lpr, pk.prt, x, y, z, lt, val,
                                          mo, Rp, a, ilig);
      }
      if (active && lead) P.ev_sur[sv] = sv_s;
      return;
    }
  }
  if (P.mode != MODE_MC) {
    place(mo);
    gd_energy E;
    pose_energy<APL, TILE>(C, S, lpr, sp, true, x, y, z, ox, oy, oz, lt, q, li, val, mo, E);
    if (active && lead) {
      if (P.mode == MODE_EVAL) {
        P.ev_out[sv] = E;
      } else {
        P.e_old[sv] = E;
        P.best_e[sv] = E.e[GD_MAXWEI - 1];
#pragma unroll
        for (int i = 0; i < 6; ++i) P.best_mm[(size_t)sv * 6 + i] = mo[i];
        if (P.rec) {
          const int rc = P.rec_cnt[sv];
          if (rc < P.rec_cap) {
            gd_record &r = P.rec[(size_t)sv * P.rec_cap + rc];
            r.replica.idx_rep = flat;
            r.replica.idx_prt = pk.prt;
            r.replica.idx_tmp = P.idx_tmp[sv];
            r.replica.idx_lig = ilig;
            r.energy = E;
            for (int i = 0; i < 6; ++i) r.movematrix[i] = mo[i];
            r.step = step0;
            P.rec_cnt[sv] = rc + 1;
          }
        }
      }
    }
    return;
  }

  // ---- MC steps ------------------------------------------------------------
  const int cid = C.cid;
  const int nc = P.nc;
  float mb = C.minus_beta[S.xi[sl]];
  int acs = 0;
  int rc = P.rec ? P.rec_cnt[sv] : 0;
  float best = P.best_e[sv];
  auto track_best = [&]() {
    if (eo.e[GD_MAXWEI - 1] < best) {
      best = eo.e[GD_MAXWEI - 1];
      if (active && lead)
#pragma unroll
        for (int i = 0; i < 6; ++i) P.best_mm[(size_t)sv * 6 + i] = mo[i];
    }
  };
  const bool wave_active = __ballot(active) != 0;
  float scale[6];
#pragma unroll
  for (int i = 0; i < 6; ++i) scale[i] = C.move_scale[i];

  auto record = [&](const gd_energy &E, int step) {
    if (P.rec && active) {
      if (lead && rc < P.rec_cap) {
        gd_record &r = P.rec[(size_t)sv * P.rec_cap + rc];
        r.replica.idx_rep = flat;
        r.replica.idx_prt = pk.prt;
        r.replica.idx_tmp = S.xi[sl];
        r.replica.idx_lig = ilig;
        r.energy = E;
        for (int i = 0; i < 6; ++i) r.movematrix[i] = mo[i];
        r.step = step;
      }
      rc += 1;
    }
  };
  // delayed acceptance (surrogate transition): K surrogate Metropolis
  // sub-steps, then one exact test of the composite move with
  // min(1, exp(-beta [(E(y) - E(x)) - (S(y) - S(x))]))
  const int K = DA ? P.da_k : 1;
  unsigned char *da = nullptr;
  float s_old = 0.0f;
  // the split's groups share the surrogate's KDE and MCS sums for complexes
  // with large KDE or MCS sets (one exchange per surrogate evaluation, so
  // every thread evaluates it together)
  const bool share = SHARED && sp.ns > 1 && C.da_share != 0;
  if constexpr (DA) {
    extern __shared__ __attribute__((aligned(16))) unsigned char lds_da[];
    da = lds_da + P.off_da + (threadIdx.x >> 6) * 6208;
    if (wave_active) place(mo);
    if constexpr (SHARED) {
      if (share)
        s_old = sur_total<APL, DAM, true>(P, C, S, sp, wave_active, da, lpr, pk.prt, x, y, z, lt,
                                          val, mo, Rp, a, ilig);
    }
    if (!share && wave_active)
      s_old = sur_total<APL, DAM, false>(P, C, S, sp, true, da, lpr, pk.prt, x, y, z, lt, val, mo,
                                         Rp, a, ilig);
  }

  for (int s = step0; s < step0 + P.n_steps; s += K) {
    if constexpr (!DA) {
      float u[8], mn[6];
      if (wave_active) {
        float dm[6];
        mc_uniforms(P.k0, P.k1, cid, flat, s, u);
        move_offsets(u, scale, C.move_kind, dm);
        apply_move(mo, dm, mn);
        place(mn);
      }
      gd_energy E;
      pose_energy<APL, TILE>(C, S, lpr, sp, wave_active, x, y, z, ox, oy, oz, lt, q, li, val, mn, E);
      if (wave_active && metropolis(E.e[GD_MAXWEI - 1], eo.e[GD_MAXWEI - 1], mb, u[6])) {
#pragma unroll
        for (int i = 0; i < 6; ++i) mo[i] = mn[i];
        eo = E;
        acs += 1;
        record(E, s);
        track_best();
      }
    } else {
      float yv[6];
#pragma unroll
      for (int i = 0; i < 6; ++i) yv[i] = mo[i];
      float s_y = s_old, u7 = 1.0f;
      bool moved = false;
      if (!share) {
        for (int j = 0; wave_active && j < K; ++j) {
          float u[8], dm[6], mn[6];
          mc_uniforms(P.k0, P.k1, cid, flat, s + j, u);
          move_offsets(u, scale, C.move_kind, dm);
          apply_move(yv, dm, mn);
          place(mn);
          const float s_n = sur_total<APL, DAM, false>(P, C, S, sp, true, da, lpr, pk.prt, x, y,
                                                       z, lt, val, mn, Rp, a, ilig);
          if (metropolis(s_n, s_y, mb, u[6])) {
#pragma unroll
            for (int i = 0; i < 6; ++i) yv[i] = mn[i];
            s_y = s_n;
            moved = true;
          }
          u7 = u[7];
        }
      } else if constexpr (SHARED) {
        // every wave runs the K evaluations: their exchanges are workgroup
        // barriers
        for (int j = 0; j < K; ++j) {
          float u[8], dm[6], mn[6];
          if (wave_active) {
            mc_uniforms(P.k0, P.k1, cid, flat, s + j, u);
            move_offsets(u, scale, C.move_kind, dm);
            apply_move(yv, dm, mn);
            place(mn);
          }
          const float s_n = sur_total<APL, DAM, true>(P, C, S, sp, wave_active, da, lpr, pk.prt,
                                                      x, y, z, lt, val, mn, Rp, a, ilig);
          if (wave_active) {
            if (metropolis(s_n, s_y, mb, u[6])) {
#pragma unroll
              for (int i = 0; i < 6; ++i) yv[i] = mn[i];
              s_y = s_n;
              moved = true;
            }
            u7 = u[7];
          }
        }
      }
      const bool eval = wave_active && __ballot(moved) != 0;
      if (eval) place(yv);
      gd_energy E;
      pose_energy<APL, TILE>(C, S, lpr, sp, eval, x, y, z, ox, oy, oz, lt, q, li, val, yv, E);
      if (eval) {
        const float d1 = E.e[GD_MAXWEI - 1] - eo.e[GD_MAXWEI - 1];
        const float d2 = s_y - s_old;
        const float delta = d1 - d2;
        const float prod = delta * mb;
        if (moved && u7 < decision_expf(prod)) {
#pragma unroll
          for (int i = 0; i < 6; ++i) mo[i] = yv[i];
          eo = E;
          s_old = s_y;
          acs += 1;
          record(E, s + K - 1);
          track_best();
        }
      }
    }
    if (nc > 0 && ((s + K) % nc) == 0) {
      if (active && lead) S.xe[slot] = eo.e[GD_MAXWEI - 1];
      __syncthreads();
      if (tid < pk.n_groups) {
        const int n_tmp = C.n_tmp;
        const int seg = (s + K) / nc - 1;
        const int mode_t = (seg % 2 == 0) ? 1 : 0;
        const int f0 = P.rep_flat[P.slots[pk.slot_off + tid * n_tmp]];
        const int group = (f0 / (n_tmp * n_lig)) * n_lig + (f0 % n_lig);
        const int b = tid * n_tmp;
        exchange_ladder(n_tmp, mode_t, S.xe + b, C.minus_beta, S.xi + b, S.xa + b, S.xo + b,
                        [&](int pair) {
                          return exchange_uniform(P.k0, P.k1, cid, group, seg, pair);
                        });
      }
      __syncthreads();
      mb = C.minus_beta[S.xi[sl]];
    }
  }

  if (active && lead) {
#pragma unroll
    for (int i = 0; i < 6; ++i) P.mm[(size_t)sv * 6 + i] = mo[i];
    P.e_old[sv] = eo;
    P.best_e[sv] = best;
    P.idx_tmp[sv] = S.xi[slot];
    P.acs_mc[sv] += acs;
    P.acs_x[sv] = S.xa[slot];
    if (P.rec) P.rec_cnt[sv] = rc < P.rec_cap ? rc : P.rec_cap;
  }
}

template <int APL, int DA, bool TILE>
__device__ __forceinline__ void mc_kernel_body(const KParams &P, int idx, int step0) {
  extern __shared__ __attribute__((aligned(16))) unsigned char lds[];
  const Pack pk = P.packs[P.pack_base + idx];
  const DevComplex &C = P.cxs[pk.cx];
  Lds S;
  S.xyze = reinterpret_cast<const float4 *>(lds);
  S.hpp = reinterpret_cast<const float *>(lds + P.off_hpp);
  S.psp = reinterpret_cast<const int *>(lds + P.off_psp);
  S.par = reinterpret_cast<const float *>(lds + P.off_par);
  S.kde = reinterpret_cast<const float4 *>(lds + P.off_kde);
  S.kr = reinterpret_cast<const int2 *>(lds + P.off_kr);
  S.ch = reinterpret_cast<const int4 *>(lds + P.off_ch);
  S.xe = reinterpret_cast<float *>(lds + P.off_xe);
  S.xi = reinterpret_cast<int *>(lds + P.off_xi);
  S.xa = reinterpret_cast<int *>(lds + P.off_xa);
  S.xo = reinterpret_cast<int *>(lds + P.off_xo);
  S.xp = reinterpret_cast<float *>(lds + P.off_xp);
  const int pnp = C.pnp;
  const float4 *src = C.prt_xyze + (size_t)pk.prt * pnp;
  S.gx = src;
  S.ghpp = C.prt_hpp;
  S.gpsp = C.prt_psp;
  S.tx = reinterpret_cast<float4 *>(lds);
  S.th = reinterpret_cast<float *>(lds + P.off_hpp);
  S.tp = reinterpret_cast<int *>(lds + P.off_psp);
  S.tile = P.tile;
  if constexpr (TILE) S.kde = C.kde;  // KDE points are read from HBM

  // stage the pack's receptor conformer (unless it streams in tiles) and tables
  const int nt = blockDim.x;
  {
    float4 *xyze = reinterpret_cast<float4 *>(lds);
    float *hpp = reinterpret_cast<float *>(lds + P.off_hpp);
    int *psp = reinterpret_cast<int *>(lds + P.off_psp);
    float *par = reinterpret_cast<float *>(lds + P.off_par);
    float4 *kde = reinterpret_cast<float4 *>(lds + P.off_kde);
    int2 *kr = reinterpret_cast<int2 *>(lds + P.off_kr);
    int4 *ch = reinterpret_cast<int4 *>(lds + P.off_ch);
    if constexpr (!TILE) {
      for (int i = threadIdx.x; i < pnp; i += nt) {
        xyze[i] = src[i];
        hpp[i] = 0.5f * C.prt_hpp[i];
        psp[i] = C.prt_psp[i];
      }
      for (int i = threadIdx.x; i < C.pnk; i += nt) kde[i] = C.kde[i];
    }
    for (int i = threadIdx.x; i < PAR_FLOATS; i += nt) par[i] = C.params[i];
    for (int i = threadIdx.x; i < NLT; i += nt) kr[i] = C.kde_range[i];
    for (int i = threadIdx.x; i < C.n_chunks; i += nt) ch[i] = C.chunks[i];
    if (P.mode == MODE_MC) {
      for (int i = threadIdx.x; i < pk.n_slots; i += nt) {
        const int g = P.slots[pk.slot_off + i];
        S.xi[i] = P.idx_tmp[g];
        S.xa[i] = P.acs_x[g];
      }
    }
  }
  __syncthreads();

  run_pack<APL, DA, TILE>(P, pk, C, S, C.lpr, step0);
}

// MC kernel for packs whose complexes place APL ligand atoms on each lane:
// plain MC, start-pose evaluation and pose evaluation (DA = false), or
// delayed acceptance with surrogate sub-steps (DA = true). One instance per
// APL keeps each variant's register allocation to what it needs.
template <int APL, int DA, bool TILE>
__global__ void __launch_bounds__(512) gd_mc_kernel(KParams P) {
  const int step0 = P.step_dev ? *P.step_dev : P.step0;
  // persistent launches: each workgroup takes packs in LPT order from a
  // counter until none is left; otherwise one pack per workgroup
  extern __shared__ __attribute__((aligned(16))) unsigned char lds_q[];
  int *next = reinterpret_cast<int *>(lds_q + P.off_q);
  const bool persistent = P.queue != nullptr;
  int idx = blockIdx.x;
  for (;;) {
    if (persistent) {
      if (threadIdx.x == 0) *next = atomicAdd(P.queue, 1);
      __syncthreads();
      idx = *next;
      __syncthreads();  // every thread has read the index; the previous pack's LDS is free
      if (idx >= P.n_queue) break;
    }
    mc_kernel_body<APL, DA, TILE>(P, idx, step0);
    if (!persistent) break;
  }
}

}  // namespace gd
