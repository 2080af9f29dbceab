/*
 * gdo.h -- CPU oracle for the GeauxDock replica-exchange MC hot path.
 *
 * TEST INFRASTRUCTURE ONLY. This code is the checker: tests/, the smoke check
 * and the benchmark's cpu_baseline leg may call it; the engine never does.
 *
 * Contents
 *   - Philox4x32-10 counter-based generator (Salmon, Moraes, Dror, Shaw,
 *     "Parallel random numbers: as easy as 1, 2, 3", SC'11), the stream layout
 *     shared with the engine, uniform (0,1] and Box-Muller normals.
 *   - gdo_expf: single-precision exp used by every Metropolis / exchange
 *     decision (Cody-Waite reduction, degree-7 Taylor polynomial, explicit fma).
 *   - FP64 per-term energy of one pose (the parity oracle) with per-term
 *     magnitudes sum|contribution| for conditioned error bounds.
 *   - FP32 CPU MC path (Move, RMSD, MCC, CalcEnergy, Accept, record,
 *     ExchangeReplicas), OpenMP over replicas: the CPU baseline.
 *
 * Reference files followed are cited per function in gdo.c.
 */
#ifndef GDO_H
#define GDO_H

#include <stdint.h>
#include "../include/geauxdock.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- random numbers ---------------------------------------------------- */
void gdo_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
float gdo_uniform(uint32_t bits); /* ((bits >> 8) + 1) * 2^-24, in (0, 1] */
/* Eight uniforms of MC step `step` of replica `flat` of complex `cid`. */
void gdo_mc_uniforms(uint64_t seed, int cid, int flat, int step, float u[8]);
/* Uniform of exchange pair `pair` of ladder `group` at segment `seg`. */
float gdo_exchange_uniform(uint64_t seed, int cid, int group, int seg, int pair);
/* Six move offsets for one MC step from u[0..5]. */
void gdo_move_offsets(const float u[8], const float scale[6], int move_kind, float dm[6]);

/* ---- decisions ----------------------------------------------------------- */
float gdo_expf(float x);
int gdo_accept(float e_new, float e_old, float minus_beta, float u);
/* One exchange sweep of one ladder; etot/idx_tmp/acs by slot, mb by
 * temperature, u by pair. */
void gdo_exchange_group(int n_tmp, int mode_t, const float *etot, const float *mb,
                        const float *u, int32_t *idx_tmp, int32_t *acs);
/* The same sweep, writing each attempted pair's margin delta - ln U to
 * margin[pair] when margin is not NULL. */
void gdo_exchange_group_log(int n_tmp, int mode_t, const float *etot, const float *mb,
                            const float *u, int32_t *idx_tmp, int32_t *acs, float *margin);

/* ---- FP64 energy oracle ------------------------------------------------- */
typedef struct gdo_energy64 {
  double e[GD_MAXWEI];   /* normalized terms and total, as gd_energy */
  double raw[GD_MAXWEI]; /* terms before a*e+b (after /lna, log, ...) */
  double mag[GD_MAXWEI]; /* |a| * sum|contribution| + |b| per term of e (condition scale) */
  double cms, rmsd;
  int64_t tp, fn, fp, tn;
  double min_margin;     /* min |d - pmf0| over all pairs (contact boundary) */
} gdo_energy64;

/* Reference contact map of replica 0's start pose (InitRefMatrix_d,
 * src/kernel_cuda_l2_calcmcc.cu:57-83), [lna*pnp] bytes, FP64 distances. */
void gdo_refmap_f64(const gd_complex_desc *cx, uint8_t *ref);
void gdo_energy_f64(const gd_complex_desc *cx, const uint8_t *ref, int iprt, int ilig,
                    const float mm[6], gdo_energy64 *out);
/* gdo_energy_f64 of n poses (pose i: protein conformer iprt[i], ligand
 * conformer ilig[i], movematrix mm[6i..6i+5]), OpenMP over poses. */
void gdo_energy_f64_batch(const gd_complex_desc *cx, const uint8_t *ref, int n, const int *iprt,
                          const int *ilig, const float *mm, gdo_energy64 *out, int n_threads);
/* Ligand coordinates of a pose in FP64 (Move_d, src/kernel_cuda_l2_move.cu:21-80). */
void gdo_place_f64(const gd_complex_desc *cx, int ilig, const float mm[6], double *x,
                   double *y, double *z);

/* ---- FP32 CPU path ------------------------------------------------------ */
void gdo_refmap_f32(const gd_complex_desc *cx, uint8_t *ref);
void gdo_energy_f32(const gd_complex_desc *cx, const uint8_t *ref, int iprt, int ilig,
                    const float mm[6], gd_energy *out);

typedef struct gdo_mc_io {
  float *mm;          /* [n_rep*6] movematrix_old */
  gd_energy *e_old;   /* [n_rep] */
  int32_t *idx_tmp;   /* [n_rep] */
  int32_t *acs_mc;    /* [n_rep] */
  int32_t *acs_x;     /* [n_rep] */
  gd_record *rec;     /* [n_rep*rec_cap] or NULL */
  int64_t rec_cap;    /* records per replica */
  int64_t *rec_cnt;   /* [n_rep] */
  /* Decision margins, optional (NULL = not logged), in FP64 of the FP32
   * operands: positive accepts. For steps s in [log_step0, log_step0+log_steps):
   *   mlog[flat*log_steps + s-log_step0]  Metropolis test of step s,
   *     minus_beta*(E_new-E_old) - ln u6 (delayed acceptance: the surrogate
   *     sub-step test of step s);
   *   clog[same index]  delayed acceptance only: the composite test of the
   *     group ending at step s, minus_beta*(dE-dS) - ln u7 (NaN elsewhere);
   *   xlog[(seg*n_ladders + g)*(GD_MAX_NTMP/2) + pair]  exchange pair `pair`
   *     of ladder g (p*n_lig + l) after segment seg < log_segs,
   *     (mb1-mb2)(E2-E1) - ln U (NaN when the pair is not attempted). */
  float *mlog, *clog, *xlog;
  int log_step0, log_steps, log_segs;
} gdo_mc_io;

/* Runs steps [step0, step0+n_steps). init=1 first evaluates the start poses
 * and records them at step0 (MonteCarlo_Init_d). nc = steps per exchange
 * (0 = none). Returns 0 on success. */
int gdo_mc_run(const gd_complex_desc *cx, uint64_t seed, int cid, int init, int step0,
               int n_steps, int nc, gdo_mc_io *io, int n_threads);

/* Delayed-acceptance MC whose surrogate is the exact cheap-term total
 * (KDE + MCS + pocket distance, weighted): k_sub surrogate Metropolis
 * sub-steps then one exact composite test. Start poses must already be
 * evaluated (e_old). step0, n_steps and nc are multiples of k_sub. */
int gdo_mc_run_da(const gd_complex_desc *cx, uint64_t seed, int cid, int step0, int n_steps,
                  int nc, int k_sub, gdo_mc_io *io, int n_threads);

#ifdef __cplusplus
}
#endif

/* post-MC clustering (post.c) */
void gdo_km_init(int n, int k, int group, int level, uint64_t seed, int *out);
int gdo_kmeans(const float *f, int n, int k, float threshold, int max_pass, int group, int level,
               uint64_t seed, int *memb, int *medoid, int *size);
int gdo_post_mc(const gd_record *recs, int64_t n, int k_rep, int k_final, uint64_t seed,
                int64_t *medoid_rec, int *medoid_size);

#endif
