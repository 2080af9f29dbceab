/*
 * gdo.c -- CPU oracle of the GeauxDock MC hot path. TEST INFRASTRUCTURE ONLY:
 * called by tests/, the smoke check and the benchmark's cpu_baseline leg,
 * never by the engine.
 *
 * Reference (GeauxDock, geaux-computational-bio/geaxdock_gpu) files followed:
 *   energy terms, order, normalization  src/kernel_cuda_l2_calcenergy.cu:19-360
 *   weighted total                      src/kernel_cuda_l3_combineenergy.cu:23-37
 *   rigid move, Euler rotation          src/kernel_cuda_l2_move.cu:21-80
 *   Metropolis test                     src/kernel_cuda_l2_accept.cu:16-42
 *   RMSD (by its definition, see Q4)    src/kernel_cuda_l2_calcrmsd.cu:1-36
 *   contact map and MCC                 src/kernel_cuda_l2_calcmcc.cu:57-83,148-213
 *   MC step order, record               src/kernel_cuda_l1_montecarlo.cu:72-128,
 *                                       src/kernel_cuda_l3_util.cu:73-108
 *   temperature exchange                src/kernel_cuda_l1_exchangereplicas.cu:29-78
 *   replica layout                      src/util.C:842-861
 *   serial loop / OpenMP shape only     src/unused/cpu_mic/montecarlo_cpu.C:28-58
 */
#include "gdo.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* Philox4x32-10 (Salmon et al., SC'11): 10 rounds of a 4x32 multiply-xor   */
/* bijection; the key is bumped by the Weyl constants between rounds.       */
/* ------------------------------------------------------------------------ */
#define PHILOX_M0 0xD2511F53u
#define PHILOX_M1 0xCD9E8D57u
#define PHILOX_W0 0x9E3779B9u
#define PHILOX_W1 0xBB67AE85u

void gdo_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int r = 0; r < 10; ++r) {
    if (r > 0) {
      k0 += PHILOX_W0;
      k1 += PHILOX_W1;
    }
    const uint64_t p0 = (uint64_t)PHILOX_M0 * c0;
    const uint64_t p1 = (uint64_t)PHILOX_M1 * c2;
    const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    c0 = hi1 ^ c1 ^ k0;
    c1 = lo1;
    c2 = hi0 ^ c3 ^ k1;
    c3 = lo0;
  }
  out[0] = c0;
  out[1] = c1;
  out[2] = c2;
  out[3] = c3;
}

float gdo_uniform(uint32_t bits) { return (float)((bits >> 8) + 1u) * 5.9604644775390625e-08f; }

void gdo_mc_uniforms(uint64_t seed, int cid, int flat, int step, float u[8]) {
  const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  for (int b = 0; b < 2; ++b) {
    const uint32_t ctr[4] = {(uint32_t)step, (uint32_t)flat, (uint32_t)cid, (uint32_t)b};
    uint32_t o[4];
    gdo_philox4x32_10(ctr, key, o);
    for (int i = 0; i < 4; ++i) u[4 * b + i] = gdo_uniform(o[i]);
  }
}

float gdo_exchange_uniform(uint64_t seed, int cid, int group, int seg, int pair) {
  const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  const uint32_t ctr[4] = {(uint32_t)seg, (uint32_t)group, (uint32_t)cid, 0x100u + (uint32_t)pair};
  uint32_t o[4];
  gdo_philox4x32_10(ctr, key, o);
  return gdo_uniform(o[0]);
}

void gdo_move_offsets(const float u[8], const float scale[6], int move_kind, float dm[6]) {
  if (move_kind == GD_MOVE_GAUSSIAN) {
    /* Box-Muller on the pairs (u0,u1), (u2,u3), (u4,u5). */
    for (int k = 0; k < 3; ++k) {
      const float r = sqrtf(-2.0f * logf(u[2 * k]));
      const float th = 6.28318530717958647692f * u[2 * k + 1];
      dm[2 * k] = scale[2 * k] * (r * cosf(th));
      dm[2 * k + 1] = scale[2 * k + 1] * (r * sinf(th));
    }
  } else {
    /* move_scale * (2U - 1), src/kernel_cuda_l1_montecarlo.cu:100 */
    for (int i = 0; i < 6; ++i) dm[i] = scale[i] * (2.0f * u[i] - 1.0f);
  }
}

/* ------------------------------------------------------------------------ */
/* exp for decisions. x = n ln2 + r with n = rint(x log2 e) and the         */
/* Cody-Waite two-constant reduction (ln2 = LN2_HI + LN2_LO, LN2_HI with 9  */
/* trailing zero bits), then e^r by its Taylor series to degree 7 in Horner */
/* form (|r| <= ln2/2, truncation < 6e-9 relative), then scaling by 2^n.     */
/* Arguments above 88 give +inf and below -87 give 0 so results never       */
/* leave the normal range.                                                   */
/* ------------------------------------------------------------------------ */
float gdo_expf(float x) {
  if (x != x) return x;
  if (x > 88.0f) return INFINITY;
  if (x < -87.0f) return 0.0f;
  const float n = rintf(x * 1.44269502162933349609375f);
  float r = fmaf(-n, 0.693145751953125f, x);
  r = fmaf(-n, 1.428606765330187045e-06f, r);
  float p = 1.98412698412698412698e-04f; /* 1/7! */
  p = fmaf(p, r, 1.38888888888888888889e-03f); /* 1/6! */
  p = fmaf(p, r, 8.33333333333333333333e-03f); /* 1/5! */
  p = fmaf(p, r, 4.16666666666666666667e-02f); /* 1/4! */
  p = fmaf(p, r, 1.66666666666666666667e-01f); /* 1/3! */
  p = fmaf(p, r, 0.5f);
  p = fmaf(p, r, 1.0f);
  p = fmaf(p, r, 1.0f);
  return ldexpf(p, (int)n);
}

/* Accept_d: is_accept = U < exp(dE * minus_beta), src/kernel_cuda_l2_accept.cu:23 */
int gdo_accept(float e_new, float e_old, float minus_beta, float u) {
  const float de = e_new - e_old;
  const float prod = de * minus_beta;
  return u < gdo_expf(prod);
}

/* ExchangeReplicas_d for one (prt, lig) ladder,
 * src/kernel_cuda_l1_exchangereplicas.cu:33-78. */
void gdo_exchange_group(int n_tmp, int mode_t, const float *etot, const float *mb,
                        const float *u, int32_t *idx_tmp, int32_t *acs) {
  gdo_exchange_group_log(n_tmp, mode_t, etot, mb, u, idx_tmp, acs, NULL);
}

void gdo_exchange_group_log(int n_tmp, int mode_t, const float *etot, const float *mb,
                            const float *u, int32_t *idx_tmp, int32_t *acs, float *margin) {
  int temps[GD_MAX_NTMP], order[GD_MAX_NTMP];
  float en[GD_MAX_NTMP];
  for (int t = 0; t < n_tmp; ++t) {
    temps[t] = idx_tmp[t];
    en[t] = etot[t];
    order[temps[t]] = t;
  }
  if (mode_t < 2) {
    const int maxpair = n_tmp / 2 - ((n_tmp % 2 == 0) && (mode_t == 1));
    for (int pair = 0; pair < maxpair; ++pair) {
      const int i1 = (pair << 1) + mode_t;
      const int i2 = i1 + 1;
      const int o1 = order[i1];
      const int o2 = order[i2];
      const int tt = temps[o1];
      const float e1 = en[o1], e2 = en[o2];
      const float mb1 = mb[temps[o1]], mb2 = mb[temps[o2]];
      const float d1 = mb1 - mb2;
      const float d2 = e2 - e1;
      const float delta = d1 * d2;
      const float prob = gdo_expf(delta);
      if (margin) margin[pair] = (float)((double)delta - log((double)u[pair]));
      if (prob > u[pair]) {
        temps[o1] = temps[o2];
        temps[o2] = tt;
        acs[o1] += 1;
        acs[o2] += 1;
      }
    }
  }
  for (int t = 0; t < n_tmp; ++t) idx_tmp[t] = temps[t];
}

/* ------------------------------------------------------------------------ */
/* FP64 oracle                                                               */
/* ------------------------------------------------------------------------ */
void gdo_place_f64(const gd_complex_desc *cx, int ilig, const float mm[6], double *x,
                   double *y, double *z) {
  /* Euler rotation of src/kernel_cuda_l2_move.cu:43-51. */
  const double s1 = sin((double)mm[3]), c1 = cos((double)mm[3]);
  const double s2 = sin((double)mm[4]), c2 = cos((double)mm[4]);
  const double s3 = sin((double)mm[5]), c3 = cos((double)mm[5]);
  const double r00 = c1 * c2, r01 = c1 * s2 * s3 - c3 * s1, r02 = s1 * s3 + c1 * c3 * s2;
  const double r10 = c2 * s1, r11 = c1 * c3 + s1 * s2 * s3, r12 = c3 * s1 * s2 - c1 * s3;
  const double r20 = -s2, r21 = c2 * s3, r22 = c2 * c3;
  const int lna = cx->lna;
  const float *ox = cx->lig_x + (size_t)ilig * lna;
  const float *oy = cx->lig_y + (size_t)ilig * lna;
  const float *oz = cx->lig_z + (size_t)ilig * lna;
  const double tx = (double)mm[0] + cx->pocket_center[0];
  const double ty = (double)mm[1] + cx->pocket_center[1];
  const double tz = (double)mm[2] + cx->pocket_center[2];
  for (int l = 0; l < lna; ++l) {
    x[l] = r00 * ox[l] + r01 * oy[l] + r02 * oz[l] + tx;
    y[l] = r10 * ox[l] + r11 * oy[l] + r12 * oz[l] + ty;
    z[l] = r20 * ox[l] + r21 * oy[l] + r22 * oz[l] + tz;
  }
}

void gdo_refmap_f64(const gd_complex_desc *cx, uint8_t *ref) {
  const int lna = cx->lna, pnp = cx->pnp;
  const float zero[6] = {0, 0, 0, 0, 0, 0};
  double *x = (double *)malloc(sizeof(double) * 3 * lna);
  double *y = x + lna, *z = y + lna;
  gdo_place_f64(cx, 0, zero, x, y, z);
  const gd_enepara *ep = cx->enepara;
  for (int l = 0; l < lna; ++l) {
    const int lt = cx->lig_t[l];
    for (int p = 0; p < pnp; ++p) {
      const double dx = x[l] - cx->prt_x[p], dy = y[l] - cx->prt_y[p], dz = z[l] - cx->prt_z[p];
      const double d = sqrt(dx * dx + dy * dy + dz * dz);
      ref[(size_t)l * pnp + p] = d <= (double)ep->pmf0[lt][cx->prt_t[p]];
    }
  }
  free(x);
}

static void finalize64(const gd_complex_desc *cx, const double raw_in[9], const double mag_in[9],
                       gdo_energy64 *out) {
  const gd_enepara *ep = cx->enepara;
  double total = 0.0;
  for (int i = 0; i < 9; ++i) {
    out->raw[i] = raw_in[i];
    /* magnitude of the normalized term a*e+b */
    out->mag[i] = fabs((double)ep->a_para[i]) * mag_in[i] + fabs((double)ep->b_para[i]);
    out->e[i] = (double)ep->a_para[i] * raw_in[i] + (double)ep->b_para[i];
    total += (double)ep->w[i] * out->e[i];
  }
  out->e[9] = total;
  out->raw[9] = total;
  double m = 0.0;
  for (int i = 0; i < 9; ++i)
    m += fabs((double)ep->w[i]) * (fabs((double)ep->a_para[i]) * mag_in[i] + fabs((double)ep->b_para[i]));
  out->mag[9] = m;
}

void gdo_energy_f64(const gd_complex_desc *cx, const uint8_t *ref, int iprt, int ilig,
                    const float mm[6], gdo_energy64 *out) {
  const int lna = cx->lna, pnp = cx->pnp;
  const gd_enepara *ep = cx->enepara;
  double *x = (double *)malloc(sizeof(double) * 3 * lna);
  double *y = x + lna, *z = y + lna;
  gdo_place_f64(cx, ilig, mm, x, y, z);
  const float *px = cx->prt_x + (size_t)iprt * pnp;
  const float *py = cx->prt_y + (size_t)iprt * pnp;
  const float *pz = cx->prt_z + (size_t)iprt * pnp;
  const double sqrt_2_pi_m1 = -1.0 / sqrt(2.0 * 3.14159265358979323846);

  double evdw = 0, eele = 0, epmf = 0, epsp = 0, ehdb = 0, ehpc = 0;
  double mvdw = 0, mele = 0, mpmf = 0, mpsp = 0, mhdb = 0, mhpc = 0;
  int64_t tp = 0, fn = 0, fp = 0, tn = 0;
  double min_margin = INFINITY;

  /* pair loop, src/kernel_cuda_l2_calcenergy.cu:92-208 */
  for (int l = 0; l < lna; ++l) {
    const int lt = cx->lig_t[l];
    const double ql = cx->lig_c[l];
    double hpc_l = 0.0, hpc_lm = 0.0;
    for (int p = 0; p < pnp; ++p) {
      const int pt = cx->prt_t[p];
      const double dx = x[l] - px[p], dy = y[l] - py[p], dz = z[l] - pz[p];
      const double d2 = dx * dx + dy * dy + dz * dz;
      const double d4 = d2 * d2;
      const double d = sqrt(d2);
      /* hydrophobic contact switch, :114-122 */
      if (cx->prt_c0_and_d12_or_c2[p] == 1 && d2 <= 81.0) {
        const double v = (double)cx->prt_hpp[p] *
                         (1.0 - (3.5 / 81.0 * d2 - 4.5 / 81.0 / 81.0 * d4 +
                                 2.5 / 81.0 / 81.0 / 81.0 * d4 * d2 -
                                 0.5 / 81.0 / 81.0 / 81.0 / 81.0 * d4 * d4));
        hpc_l += v;
        hpc_lm += fabs(v);
      }
      /* soft Lennard-Jones, :130-135 */
      {
        const double p1 = (double)ep->p1a[lt][pt] / (d4 * d4 * d);
        const double p2 = (double)ep->p2a[lt][pt] / (d4 * d2);
        const double p4 = p1 * (double)ep->lj0 * (1.0 + (double)ep->lj1 * d2) + 1.0;
        const double v = (p1 - p2) / p4;
        evdw += v;
        mvdw += fabs(v);
      }
      /* soft electrostatics, :140-148 */
      {
        const double s1 = (double)ep->el1 * d;
        double g1;
        if (s1 < 1.0)
          g1 = (double)ep->el0 + (double)ep->a1 * s1 * s1 + (double)ep->b1 * s1 * s1 * s1;
        else
          g1 = 1.0 / s1;
        const double v = ql * (double)cx->prt_ele[p] * g1;
        eele += v;
        mele += fabs(v);
      }
      /* contact PMF and pocket-specific potential, :157-171 */
      const double dm = d - (double)ep->pmf0[lt][pt];
      {
        const double v = (double)ep->pmf1[lt][pt] / (1.0 + exp((-0.5 * d + 6.0) * dm));
        epmf += v;
        mpmf += fabs(v);
        if (fabs(dm) < min_margin) min_margin = fabs(dm);
      }
      if (cx->prt_c[p] == 2 && dm <= 0.0) {
        const double v = (double)cx->psp[(size_t)lt * cx->n_res + cx->prt_seq3r[p]];
        epsp += v;
        mpsp += fabs(v);
      }
      /* hydrogen bond, :179-184 */
      {
        const double hdb0 = ep->hdb0[lt][pt];
        if (hdb0 > 0.1) {
          const double hdb1 = ep->hdb1[lt][pt];
          const double h3 = (d - hdb0) * hdb1;
          const double v = sqrt_2_pi_m1 * hdb1 * exp(-0.5 * h3 * h3);
          ehdb += v;
          mhdb += fabs(v);
        }
      }
      /* contact map for MCC, src/kernel_cuda_l2_calcmcc.cu:173-186 */
      if (ref) {
        const int r = ref[(size_t)l * pnp + p];
        const int c = dm <= 0.0;
        tp += (r == 1 && c);
        fn += (r == 1 && !c);
        fp += (r == 0 && c);
        tn += (r == 0 && !c);
      }
    }
    /* hydrophobic restraint per ligand atom, :200-205 */
    const double h2 = (hpc_l - (double)ep->hpl0[lt]) / (double)ep->hpl1[lt];
    const double v = 0.5 * h2 * h2 - (double)ep->hpl2[lt];
    ehpc += v;
    /* magnitude: |restraint| plus its first-order sensitivity to the sum */
    mhpc += fabs(0.5 * h2 * h2) + fabs((double)ep->hpl2[lt]) +
            fabs(h2 / (double)ep->hpl1[lt]) * hpc_lm;
  }

  /* KDE, :224-265 */
  double ekde = 0.0, mkde = 0.0;
  for (int l = 0; l < lna; ++l) {
    double s = 0.0;
    int n = 0;
    for (int k = 0; k < cx->pnk; ++k) {
      if (cx->lig_t[l] == cx->kde_t[k]) {
        const double dx = x[l] - cx->kde_x[k], dy = y[l] - cx->kde_y[k], dz = z[l] - cx->kde_z[k];
        s += exp((double)ep->kde2 * (dx * dx + dy * dy + dz * dz));
        n++;
      }
    }
    if (n != 0) ekde += s / (double)n;
  }
  ekde = ekde / (double)ep->kde3;
  mkde = fabs(ekde);

  /* MCS position restraints, :274-313 and :341 */
  double elhm = 0.0;
  for (int m = 0; m < cx->pos; ++m) {
    double s = 0.0;
    int n = 0;
    for (int l = 0; l < lna; ++l) {
      const int idx = cx->lig_n[l] + 1;
      const size_t o = (size_t)m * cx->n_mcs + idx;
      if (cx->mcs_x[o] != GD_MCS_INVALID_COORD) {
        const double dx = x[l] - cx->mcs_x[o], dy = y[l] - cx->mcs_y[o], dz = z[l] - cx->mcs_z[o];
        s += dx * dx + dy * dy + dz * dz;
        n++;
      }
    }
    if (n != 0) elhm += (double)cx->mcs_tcc[m] * sqrt(s / (double)n);
  }
  const double e7 = cx->pos > 0 ? log(elhm / (double)cx->pos) : 0.0;

  /* pocket distance, :321-326 (equals |translation|) */
  const double e8 = sqrt((double)mm[0] * mm[0] + (double)mm[1] * mm[1] + (double)mm[2] * mm[2]);

  /* RMSD against the start pose (definition of src/kernel_cuda_l2_calcrmsd.cu) */
  double rs = 0.0;
  {
    const float *ox = cx->lig_x + (size_t)ilig * lna;
    const float *oy = cx->lig_y + (size_t)ilig * lna;
    const float *oz = cx->lig_z + (size_t)ilig * lna;
    for (int l = 0; l < lna; ++l) {
      const double dx = x[l] - ((double)ox[l] + cx->pocket_center[0]);
      const double dy = y[l] - ((double)oy[l] + cx->pocket_center[1]);
      const double dz = z[l] - ((double)oz[l] + cx->pocket_center[2]);
      rs += dx * dx + dy * dy + dz * dz;
    }
  }

  const double inv = 1.0 / (double)lna;
  const double raw[9] = {evdw * inv, eele * inv, epmf * inv, epsp * inv, ehdb * inv,
                         ehpc * inv, ekde * inv, e7, e8};
  const double mag[9] = {mvdw * inv, mele * inv, mpmf * inv, mpsp * inv, mhdb * inv,
                         mhpc * inv, mkde * inv, fabs(e7), fabs(e8)};
  finalize64(cx, raw, mag, out);
  out->rmsd = sqrt(rs / (double)lna);
  out->tp = tp;
  out->fn = fn;
  out->fp = fp;
  out->tn = tn;
  {
    const double t0 = (double)tp, f0 = (double)fn, q0 = (double)fp, n0 = (double)tn;
    const double dv = sqrt((t0 + q0) * (t0 + f0) * (n0 + q0) * (n0 + f0));
    out->cms = dv != 0.0 ? (t0 * n0 - q0 * f0) / dv : 0.0;
  }
  out->min_margin = min_margin;
  free(x);
}

void gdo_energy_f64_batch(const gd_complex_desc *cx, const uint8_t *ref, int n, const int *iprt,
                          const int *ilig, const float *mm, gdo_energy64 *out, int n_threads) {
#ifdef _OPENMP
  if (n_threads > 0) omp_set_num_threads(n_threads);
#else
  (void)n_threads;
#endif
#pragma omp parallel for schedule(dynamic, 4)
  for (int i = 0; i < n; ++i) gdo_energy_f64(cx, ref, iprt[i], ilig[i], mm + 6 * (size_t)i, out + i);
}

/* ------------------------------------------------------------------------ */
/* FP32 CPU path: the reference's arithmetic in single precision.           */
/* ------------------------------------------------------------------------ */
static void place_f32(const gd_complex_desc *cx, int ilig, const float mm[6], float *x, float *y,
                      float *z) {
  const float s1 = sinf(mm[3]), c1 = cosf(mm[3]);
  const float s2 = sinf(mm[4]), c2 = cosf(mm[4]);
  const float s3 = sinf(mm[5]), c3 = cosf(mm[5]);
  const float r00 = c1 * c2, r01 = c1 * s2 * s3 - c3 * s1, r02 = s1 * s3 + c1 * c3 * s2;
  const float r10 = c2 * s1, r11 = c1 * c3 + s1 * s2 * s3, r12 = c3 * s1 * s2 - c1 * s3;
  const float r20 = -1 * s2, r21 = c2 * s3, r22 = c2 * c3;
  const int lna = cx->lna;
  const float *ox = cx->lig_x + (size_t)ilig * lna;
  const float *oy = cx->lig_y + (size_t)ilig * lna;
  const float *oz = cx->lig_z + (size_t)ilig * lna;
  const float cxx = cx->pocket_center[0], cyy = cx->pocket_center[1], czz = cx->pocket_center[2];
  for (int l = 0; l < lna; ++l) {
    x[l] = r00 * ox[l] + r01 * oy[l] + r02 * oz[l] + mm[0] + cxx;
    y[l] = r10 * ox[l] + r11 * oy[l] + r12 * oz[l] + mm[1] + cyy;
    z[l] = r20 * ox[l] + r21 * oy[l] + r22 * oz[l] + mm[2] + czz;
  }
}

void gdo_refmap_f32(const gd_complex_desc *cx, uint8_t *ref) {
  const int lna = cx->lna, pnp = cx->pnp;
  const float zero[6] = {0, 0, 0, 0, 0, 0};
  float x[GD_MAX_LNA], y[GD_MAX_LNA], z[GD_MAX_LNA];
  place_f32(cx, 0, zero, x, y, z);
  for (int l = 0; l < lna; ++l) {
    const int lt = cx->lig_t[l];
    for (int p = 0; p < pnp; ++p) {
      const float dx = x[l] - cx->prt_x[p], dy = y[l] - cx->prt_y[p], dz = z[l] - cx->prt_z[p];
      const float d = sqrtf(dx * dx + dy * dy + dz * dz);
      ref[(size_t)l * pnp + p] = d <= cx->enepara->pmf0[lt][cx->prt_t[p]];
    }
  }
}

void gdo_energy_f32(const gd_complex_desc *cx, const uint8_t *ref, int iprt, int ilig,
                    const float mm[6], gd_energy *out) {
  const int lna = cx->lna, pnp = cx->pnp;
  const gd_enepara *ep = cx->enepara;
  float x[GD_MAX_LNA], y[GD_MAX_LNA], z[GD_MAX_LNA];
  place_f32(cx, ilig, mm, x, y, z);
  const float *px = cx->prt_x + (size_t)iprt * pnp;
  const float *py = cx->prt_y + (size_t)iprt * pnp;
  const float *pz = cx->prt_z + (size_t)iprt * pnp;
  const float sqrt_2_pi_m1 = -1.0f / sqrtf(2.0f * 3.1415926535f);
  const float lj0 = ep->lj0, lj1 = ep->lj1, el0 = ep->el0, el1 = ep->el1, a1 = ep->a1, b1 = ep->b1;

  float evdw = 0, eele = 0, epmf = 0, epsp = 0, ehdb = 0, ehpc = 0;
  int tp = 0, fn = 0, fp = 0, tn = 0;
  for (int l = 0; l < lna; ++l) {
    const int lt = cx->lig_t[l];
    const float xl = x[l], yl = y[l], zl = z[l], ql = cx->lig_c[l];
    const float *p1a = ep->p1a[lt], *p2a = ep->p2a[lt], *pmf0 = ep->pmf0[lt], *pmf1 = ep->pmf1[lt];
    const float *hdb0t = ep->hdb0[lt], *hdb1t = ep->hdb1[lt];
    const uint8_t *refl = ref ? ref + (size_t)l * pnp : NULL;
    float vdw = 0, ele = 0, pmf = 0, psp = 0, hdb = 0, hpc = 0;
    for (int p = 0; p < pnp; ++p) {
      const int pt = cx->prt_t[p];
      const float dx = xl - px[p], dy = yl - py[p], dz = zl - pz[p];
      const float d2 = dx * dx + dy * dy + dz * dz;
      const float d4 = d2 * d2;
      const float d = sqrtf(d2);
      if (cx->prt_c0_and_d12_or_c2[p] == 1 && d2 <= 81.0f)
        hpc += cx->prt_hpp[p] *
               (1.0f - (3.5f / 81.0f * d2 - 4.5f / 81.0f / 81.0f * d4 +
                        2.5f / 81.0f / 81.0f / 81.0f * d4 * d2 -
                        0.5f / 81.0f / 81.0f / 81.0f / 81.0f * d4 * d4));
      const float p1 = p1a[pt] / (d4 * d4 * d);
      const float p2 = p2a[pt] / (d4 * d2);
      const float p4 = p1 * lj0 * (1.0f + lj1 * d2) + 1.0f;
      vdw += (p1 - p2) / p4;
      const float s1 = el1 * d;
      const float g1 = s1 < 1.0f ? el0 + a1 * s1 * s1 + b1 * s1 * s1 * s1 : 1.0f / s1;
      ele += ql * cx->prt_ele[p] * g1;
      const float dm = d - pmf0[pt];
      pmf += pmf1[pt] / (1.0f + expf((-0.5f * d + 6.0f) * dm));
      if (cx->prt_c[p] == 2 && dm <= 0.0f) psp += cx->psp[(size_t)lt * cx->n_res + cx->prt_seq3r[p]];
      const float h0 = hdb0t[pt];
      if (h0 > 0.1f) {
        const float h1 = hdb1t[pt];
        const float h3 = (d - h0) * h1;
        hdb += sqrt_2_pi_m1 * h1 * expf(-0.5f * h3 * h3);
      }
      if (refl) {
        const int r = refl[p];
        const int c = dm <= 0.0f;
        tp += (r == 1 && c);
        fn += (r == 1 && !c);
        fp += (r == 0 && c);
        tn += (r == 0 && !c);
      }
    }
    evdw += vdw;
    eele += ele;
    epmf += pmf;
    epsp += psp;
    ehdb += hdb;
    const float h2 = (hpc - ep->hpl0[lt]) / ep->hpl1[lt];
    ehpc += 0.5f * h2 * h2 - ep->hpl2[lt];
  }

  float ekde = 0.0f;
  for (int l = 0; l < lna; ++l) {
    float s = 0.0f;
    int n = 0;
    for (int k = 0; k < cx->pnk; ++k) {
      if (cx->lig_t[l] == cx->kde_t[k]) {
        const float dx = x[l] - cx->kde_x[k], dy = y[l] - cx->kde_y[k], dz = z[l] - cx->kde_z[k];
        s += expf(ep->kde2 * (dx * dx + dy * dy + dz * dz));
        n++;
      }
    }
    if (n != 0) ekde += s / (float)n;
  }
  ekde = ekde / ep->kde3;

  float elhm = 0.0f;
  for (int m = 0; m < cx->pos; ++m) {
    float s = 0.0f;
    int n = 0;
    for (int l = 0; l < lna; ++l) {
      const size_t o = (size_t)m * cx->n_mcs + cx->lig_n[l] + 1;
      if (cx->mcs_x[o] != GD_MCS_INVALID_COORD) {
        const float dx = x[l] - cx->mcs_x[o], dy = y[l] - cx->mcs_y[o], dz = z[l] - cx->mcs_z[o];
        s += dx * dx + dy * dy + dz * dz;
        n++;
      }
    }
    if (n != 0) elhm += cx->mcs_tcc[m] * sqrtf(s / (float)n);
  }

  float e[GD_MAXWEI];
  e[0] = evdw / lna;
  e[1] = eele / lna;
  e[2] = epmf / lna;
  e[3] = epsp / lna;
  e[4] = ehdb / lna;
  e[5] = ehpc / lna;
  e[6] = ekde / lna;
  e[7] = cx->pos > 0 ? logf(elhm / cx->pos) : 0.0f;
  {
    const float cxn = cx->pocket_center[0] + mm[0], cyn = cx->pocket_center[1] + mm[1],
                czn = cx->pocket_center[2] + mm[2];
    const float dx = cxn - cx->pocket_center[0], dy = cyn - cx->pocket_center[1],
                dz = czn - cx->pocket_center[2];
    e[8] = sqrtf(dx * dx + dy * dy + dz * dz);
  }
  float total = 0.0f;
  for (int i = 0; i < GD_MAXWEI - 1; ++i) {
    e[i] = ep->a_para[i] * e[i] + ep->b_para[i];
    out->e[i] = e[i];
    total += e[i] * ep->w[i];
  }
  out->e[9] = total;

  float rs = 0.0f;
  {
    const float *ox = cx->lig_x + (size_t)ilig * lna;
    const float *oy = cx->lig_y + (size_t)ilig * lna;
    const float *oz = cx->lig_z + (size_t)ilig * lna;
    for (int l = 0; l < lna; ++l) {
      const float dx = x[l] - (ox[l] + cx->pocket_center[0]);
      const float dy = y[l] - (oy[l] + cx->pocket_center[1]);
      const float dz = z[l] - (oz[l] + cx->pocket_center[2]);
      rs += dx * dx + dy * dy + dz * dz;
    }
  }
  out->rmsd = sqrtf(rs / lna);
  {
    const float t0 = (float)tp, f0 = (float)fn, q0 = (float)fp, n0 = (float)tn;
    const float dv = sqrtf((t0 + q0) * (t0 + f0) * (n0 + q0) * (n0 + f0));
    out->cms = dv != 0.0f ? (t0 * n0 - q0 * f0) / dv : 0.0f;
  }
}

static void put_record(gdo_mc_io *io, int flat, const gd_replica *rp, const gd_energy *e,
                       const float *mm, int step) {
  if (!io->rec) return;
  const int64_t c = io->rec_cnt[flat];
  if (c >= io->rec_cap) return;
  gd_record *r = &io->rec[(size_t)flat * io->rec_cap + c];
  r->replica = *rp;
  r->energy = *e;
  for (int i = 0; i < 6; ++i) r->movematrix[i] = mm[i];
  r->step = step;
  io->rec_cnt[flat] = c + 1;
}

/* Decision-log slot of (replica, step), or NULL when not logged. */
static float *log_slot(const gdo_mc_io *io, float *log, int flat, int step) {
  if (!log) return NULL;
  const int i = step - io->log_step0;
  if (i < 0 || i >= io->log_steps) return NULL;
  return log + (size_t)flat * io->log_steps + i;
}

/* Margin of the Metropolis test u < exp((e_new - e_old) * minus_beta), with
 * the FP32 operands the test uses: (e_new - e_old) * minus_beta - ln u. */
static float margin_of(float e_new, float e_old, float minus_beta, float u) {
  const float prod = (e_new - e_old) * minus_beta;
  return (float)((double)prod - log((double)u));
}

/* MonteCarlo_Init_d + MonteCarlo_d + ExchangeReplicas_d on the CPU. */
/* Replica exchange after segment seg for every (prt, lig) ladder. */
static void exchange_all(const gd_complex_desc *cx, uint64_t seed, int cid, int seg,
                         gdo_mc_io *io) {
  const int n_lig = cx->n_lig, n_prt = cx->n_prt, n_tmp = cx->n_tmp;
  const int mode_t = (seg % 2 == 0) ? 1 : 0;
  float mb[GD_MAX_NTMP];
  for (int t = 0; t < n_tmp; ++t) mb[t] = cx->temp[t].minus_beta;
  for (int p = 0; p < n_prt; ++p) {
    for (int l = 0; l < n_lig; ++l) {
      const int g = p * n_lig + l;
      float et[GD_MAX_NTMP], u[GD_MAX_NTMP];
      int32_t it[GD_MAX_NTMP], ac[GD_MAX_NTMP];
      for (int t = 0; t < n_tmp; ++t) {
        const int f = n_tmp * n_lig * p + n_lig * t + l;
        et[t] = io->e_old[f].e[9];
        it[t] = io->idx_tmp[f];
        ac[t] = io->acs_x[f];
      }
      for (int k = 0; k < n_tmp / 2; ++k) u[k] = gdo_exchange_uniform(seed, cid, g, seg, k);
      float *xm = NULL;
      if (io->xlog && seg >= 0 && seg < io->log_segs) {
        xm = io->xlog + ((size_t)seg * n_prt * n_lig + g) * (GD_MAX_NTMP / 2);
        for (int k = 0; k < GD_MAX_NTMP / 2; ++k) xm[k] = NAN;
      }
      gdo_exchange_group_log(n_tmp, mode_t, et, mb, u, it, ac, xm);
      for (int t = 0; t < n_tmp; ++t) {
        const int f = n_tmp * n_lig * p + n_lig * t + l;
        io->idx_tmp[f] = it[t];
        io->acs_x[f] = ac[t];
      }
    }
  }
}

int gdo_mc_run(const gd_complex_desc *cx, uint64_t seed, int cid, int init, int step0,
               int n_steps, int nc, gdo_mc_io *io, int n_threads) {
  const int n_lig = cx->n_lig, n_prt = cx->n_prt, n_tmp = cx->n_tmp;
  const int n_rep = n_lig * n_prt * n_tmp;
  if (cx->lna > GD_MAX_LNA || n_tmp > GD_MAX_NTMP) return 2;
  uint8_t *ref = (uint8_t *)malloc((size_t)cx->lna * cx->pnp);
  if (!ref) return 5;
  gdo_refmap_f32(cx, ref);
#ifdef _OPENMP
  if (n_threads > 0) omp_set_num_threads(n_threads);
#else
  (void)n_threads;
#endif

  if (init) {
#pragma omp parallel for schedule(dynamic, 1)
    for (int flat = 0; flat < n_rep; ++flat) {
      const int p = flat / (n_tmp * n_lig), l = flat % n_lig;
      gd_energy e;
      gdo_energy_f32(cx, ref, p, l, io->mm + 6 * (size_t)flat, &e);
      io->e_old[flat] = e;
      const gd_replica rp = {flat, p, io->idx_tmp[flat], l};
      put_record(io, flat, &rp, &e, io->mm + 6 * (size_t)flat, step0);
    }
  }

  int s = step0;
  const int s_end = step0 + n_steps;
  while (s < s_end) {
    int blk_end = s_end;
    if (nc > 0) {
      const int next_x = (s / nc + 1) * nc; /* exchange after step next_x-1 */
      if (next_x < blk_end) blk_end = next_x;
    }
    const int sb = s;
#pragma omp parallel for schedule(dynamic, 1)
    for (int flat = 0; flat < n_rep; ++flat) {
      const int p = flat / (n_tmp * n_lig), l = flat % n_lig;
      float *mm = io->mm + 6 * (size_t)flat;
      for (int st = sb; st < blk_end; ++st) {
        float u[8], dm[6], mn[6];
        gdo_mc_uniforms(seed, cid, flat, st, u);
        gdo_move_offsets(u, cx->move_scale, cx->move_kind, dm);
        for (int i = 0; i < 6; ++i) mn[i] = mm[i] + dm[i];
        gd_energy en;
        gdo_energy_f32(cx, ref, p, l, mn, &en);
        const float mb = cx->temp[io->idx_tmp[flat]].minus_beta;
        float *ml = log_slot(io, io->mlog, flat, st);
        if (ml) *ml = margin_of(en.e[9], io->e_old[flat].e[9], mb, u[6]);
        if (gdo_accept(en.e[9], io->e_old[flat].e[9], mb, u[6])) {
          for (int i = 0; i < 6; ++i) mm[i] = mn[i];
          io->e_old[flat] = en;
          io->acs_mc[flat] += 1;
          const gd_replica rp = {flat, p, io->idx_tmp[flat], l};
          put_record(io, flat, &rp, &en, mm, st);
        }
      }
    }
    s = blk_end;
    if (nc > 0 && s % nc == 0) exchange_all(cx, seed, cid, s / nc - 1, io);
  }
  free(ref);
  return 0;
}

/* Weighted sum of the terms that need no receptor pairs (KDE, MCS, pocket
 * distance): the exact part of the delayed-acceptance surrogate. */
static float cheap_total(const gd_complex_desc *cx, const gd_energy *e) {
  const gd_enepara *ep = cx->enepara;
  float t = 0.0f;
  for (int i = 6; i < 9; ++i) t += ep->w[i] * e->e[i];
  return t;
}

/* Delayed-acceptance MC with the surrogate reduced to its exact cheap terms
 * (surrogate scale 0): K Metropolis sub-steps on S, then the composite move
 * is accepted when U7 < exp(-beta ((E(y) - E(x)) - (S(y) - S(x)))). The
 * uniforms of sub-step j are those of step s + j; U7 is the last sub-step's
 * eighth uniform; an accepted composite move is recorded at step s + K - 1. */
int gdo_mc_run_da(const gd_complex_desc *cx, uint64_t seed, int cid, int step0, int n_steps,
                  int nc, int k_sub, gdo_mc_io *io, int n_threads) {
  const int n_lig = cx->n_lig, n_prt = cx->n_prt, n_tmp = cx->n_tmp;
  const int n_rep = n_lig * n_prt * n_tmp;
  if (cx->lna > GD_MAX_LNA || n_tmp > GD_MAX_NTMP) return 2;
  if (k_sub < 1 || n_steps % k_sub || step0 % k_sub || (nc > 0 && nc % k_sub)) return 1;
  uint8_t *ref = (uint8_t *)malloc((size_t)cx->lna * cx->pnp);
  if (!ref) return 5;
  gdo_refmap_f32(cx, ref);
#ifdef _OPENMP
  if (n_threads > 0) omp_set_num_threads(n_threads);
#else
  (void)n_threads;
#endif
  int s = step0;
  const int s_end = step0 + n_steps;
  while (s < s_end) {
    int blk_end = s_end;
    if (nc > 0) {
      const int next_x = (s / nc + 1) * nc;
      if (next_x < blk_end) blk_end = next_x;
    }
    const int sb = s;
#pragma omp parallel for schedule(dynamic, 1)
    for (int flat = 0; flat < n_rep; ++flat) {
      const int p = flat / (n_tmp * n_lig), l = flat % n_lig;
      float *mm = io->mm + 6 * (size_t)flat;
      gd_energy ex;
      gdo_energy_f32(cx, ref, p, l, mm, &ex);
      float s_old = cheap_total(cx, &ex);
      for (int st = sb; st < blk_end; st += k_sub) {
        const float mb = cx->temp[io->idx_tmp[flat]].minus_beta;
        float yv[6], s_y = s_old, u7 = 1.0f;
        int moved = 0;
        for (int i = 0; i < 6; ++i) yv[i] = mm[i];
        for (int j = 0; j < k_sub; ++j) {
          float u[8], dm[6], mn[6];
          gdo_mc_uniforms(seed, cid, flat, st + j, u);
          gdo_move_offsets(u, cx->move_scale, cx->move_kind, dm);
          for (int i = 0; i < 6; ++i) mn[i] = yv[i] + dm[i];
          gd_energy en;
          gdo_energy_f32(cx, ref, p, l, mn, &en);
          const float s_n = cheap_total(cx, &en);
          float *ml = log_slot(io, io->mlog, flat, st + j);
          if (ml) *ml = margin_of(s_n, s_y, mb, u[6]);
          float *cl = log_slot(io, io->clog, flat, st + j);
          if (cl) *cl = NAN;
          if (gdo_accept(s_n, s_y, mb, u[6])) {
            for (int i = 0; i < 6; ++i) yv[i] = mn[i];
            s_y = s_n;
            moved = 1;
          }
          u7 = u[7];
        }
        if (!moved) continue;
        gd_energy E;
        gdo_energy_f32(cx, ref, p, l, yv, &E);
        const float d1 = E.e[9] - io->e_old[flat].e[9];
        const float d2 = s_y - s_old;
        const float delta = d1 - d2;
        float *cl = log_slot(io, io->clog, flat, st + k_sub - 1);
        if (cl) *cl = (float)((double)(delta * mb) - log((double)u7));
        if (u7 < gdo_expf(delta * mb)) {
          for (int i = 0; i < 6; ++i) mm[i] = yv[i];
          io->e_old[flat] = E;
          s_old = s_y;
          io->acs_mc[flat] += 1;
          const gd_replica rp = {flat, p, io->idx_tmp[flat], l};
          put_record(io, flat, &rp, &E, mm, st + k_sub - 1);
        }
      }
    }
    s = blk_end;
    if (nc > 0 && s % nc == 0) exchange_all(cx, seed, cid, s / nc - 1, io);
  }
  free(ref);
  return 0;
}
