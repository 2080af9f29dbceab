// gd_util.hip -- small kernels beside the MC launch: pack-queue and best-pose
// resets, the reference contact map (InitRefMatrix_d), the decision and
// random-stream parity hooks, and the record gather before a device-to-host
// copy.
#include "gd_mc.h"

namespace gd {

// Zero the per-class pack counters of persistent launches.
extern "C" __global__ void gd_queue_reset_kernel(int *queue, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) queue[i] = 0;
}

// Best-so-far := current state, for every replica.
extern "C" __global__ void gd_reset_best_kernel(int n, const gd_energy *e_old, const float *mm,
                                                float *best_e, float *best_mm) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  best_e[i] = e_old[i].e[GD_MAXWEI - 1];
  for (int k = 0; k < 6; ++k) best_mm[(size_t)i * 6 + k] = mm[(size_t)i * 6 + k];
}

// Reference contact map of a complex (InitRefMatrix_d,
// src/kernel_cuda_l2_calcmcc.cu:57-83): contacts of replica 0 (ligand
// conformer 0, protein conformer 0) in its start pose, as per-atom lists of
// sorted receptor indices. One thread per ligand atom; run before any MCC.
extern "C" __global__ void gd_ref_kernel(DevComplex *cxs, int cxi, int *overflow) {
  DevComplex &C = cxs[cxi];
  const int l = blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= C.lna) return;
  const int lna = C.lna;
  const float m[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  const Rot R = euler(m[3], m[4], m[5]);
  const float ox = C.lig_orig[l], oy = C.lig_orig[lna + l], oz = C.lig_orig[2 * lna + l];
  const float x = R.r00 * ox + R.r01 * oy + R.r02 * oz + m[0] + C.pocket[0];
  const float y = R.r10 * ox + R.r11 * oy + R.r12 * oz + m[1] + C.pocket[1];
  const float z = R.r20 * ox + R.r21 * oy + R.r22 * oz + m[2] + C.pocket[2];
  const int lt = C.lig_t[l];
  int n = 0;
  for (int ci = 0; ci < C.n_chunks; ++ci) {
    const int4 ch = C.chunks[ci];
    const float pmf0 = C.params[(ch.z * NLT + lt) * NPAR + P_PMF0];
    for (int p = ch.x; p < ch.x + ch.y; ++p) {
      const float4 P = C.prt_xyze[p];
      const float d2 = pair_d2(x, y, z, P.x, P.y, P.z);
      const float d = d2 * fast_rsq(d2);
      if (d <= pmf0) {
        if (n < KMAX) C.ref[(size_t)l * KMAX + n] = make_int2(p, __float_as_int(pmf0));
        ++n;
      }
    }
  }
  if (n > KMAX) {
    atomicAdd(overflow, 1);
    n = KMAX;
  }
  C.ref_cnt[l] = n;
}

extern "C" __global__ void gd_accept_kernel(int n, const float *e_new, const float *e_old,
                                            const float *mb, const float *u, int *acc) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) acc[i] = metropolis(e_new[i], e_old[i], mb[i], u[i]) ? 1 : 0;
}

extern "C" __global__ void gd_exchange_kernel(int n_groups, int n_tmp, int mode_t,
                                              const float *etot, const float *mb,
                                              const float *u, int *idx_tmp, int *acs) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n_groups) return;
  int order[GD_MAX_NTMP];
  const float *ug = u + (size_t)g * (n_tmp / 2);
  exchange_ladder(n_tmp, mode_t, etot + (size_t)g * n_tmp, mb, idx_tmp + (size_t)g * n_tmp,
                  acs + (size_t)g * n_tmp, order, [&](int pair) { return ug[pair]; });
}

// Replica exchange of whole temperature ladders between MC launches (spread
// packing, where a ladder's replicas sit in several workgroups): the in-kernel
// exchange of run_pack with the same energies, temperature map, counters and
// uniforms. One thread per ladder: ladders[g] = (complex, ladder id within the
// complex, replica of temperature slot 0, replica stride between slots).
// The exchange follows the MC step that ends at step_end (a multiple of nc).
extern "C" __global__ void gd_ladder_exchange_kernel(const DevComplex *cxs, const int4 *ladders,
                                                     int n_ladders, int step_end, int nc,
                                                     uint32_t k0, uint32_t k1,
                                                     const gd_energy *e_old, int *idx_tmp,
                                                     int *acs_x) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n_ladders) return;
  const int4 L = ladders[g];
  const DevComplex &C = cxs[L.x];
  const int n_tmp = C.n_tmp;
  const int seg = step_end / nc - 1;
  const int mode_t = (seg % 2 == 0) ? 1 : 0;
  float et[GD_MAX_NTMP];
  int temps[GD_MAX_NTMP], acs[GD_MAX_NTMP], order[GD_MAX_NTMP];
  for (int t = 0; t < n_tmp; ++t) {
    const int r = L.z + L.w * t;
    et[t] = e_old[r].e[GD_MAXWEI - 1];
    temps[t] = idx_tmp[r];
    acs[t] = acs_x[r];
  }
  const int cid = C.cid, group = L.y;
  exchange_ladder(n_tmp, mode_t, et, C.minus_beta, temps, acs, order,
                  [&](int pair) { return exchange_uniform(k0, k1, cid, group, seg, pair); });
  for (int t = 0; t < n_tmp; ++t) {
    const int r = L.z + L.w * t;
    idx_tmp[r] = temps[t];
    acs_x[r] = acs[t];
  }
}

extern "C" __global__ void gd_philox_kernel(int n, const uint32_t *ctr, const uint32_t *key,
                                            uint32_t *out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const u32x4 r = philox4x32_10({ctr[4 * i], ctr[4 * i + 1], ctr[4 * i + 2], ctr[4 * i + 3]},
                                key[2 * i], key[2 * i + 1]);
  out[4 * i] = r.x;
  out[4 * i + 1] = r.y;
  out[4 * i + 2] = r.z;
  out[4 * i + 3] = r.w;
}

// Gathers each replica's window of records into one dense buffer
// (replica order, then step), ahead of one device-to-host copy.
extern "C" __global__ void gd_compact_kernel(int n_rep, const gd_record *rec, const int *cnt,
                                             const int64_t *off, int cap, gd_record *out) {
  const int r = blockIdx.x;
  if (r >= n_rep) return;
  const int n = cnt[r];
  const int64_t o = off[r];
  const int *src = reinterpret_cast<const int *>(rec + (size_t)r * cap);
  int *dst = reinterpret_cast<int *>(out + o);
  const int words = n * (int)(sizeof(gd_record) / 4);
  for (int i = threadIdx.x; i < words; i += blockDim.x) dst[i] = src[i];
}

}  // namespace gd
