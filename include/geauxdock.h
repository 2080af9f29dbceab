/*
 * geauxdock.h -- C ABI of the MI355X replica-exchange Monte Carlo docking engine.
 *
 * The reference (GeauxDock, geaux-computational-bio/geaxdock_gpu) exposes its
 * MC hot path through one C++ function,
 *     void Run(const Ligand*, const Protein*, const Psp*, const Kde*,
 *              const Mcs*, const EnePara*, const Temp*, const Replica*,
 *              const McPara*, McLog*, map<int, vector<LigRecordSingleStep>>&,
 *              const ComplexSize)                      (src/run.h:10-22)
 * that uploads the structures, runs the launch loop of
 * src/kernel_cuda_lancher_siglegpu.cu:5-69 and returns accepted-step records.
 * This header replaces that boundary with plain C entry points: sizes are
 * runtime values (no MAXLIG/MAXPRO arrays), every call returns a status code
 * instead of exit(), device state lives in an opaque context instead of
 * __constant__ symbols, and many complexes can share one batched launch.
 *
 * Data layouts that must match the reference's meaning keep its field names
 * and order: gd_enepara = EnePara (src/dock_soa.h:99-133), gd_temp = Temp
 * (src/dock.h:288-293), gd_replica = Replica (src/dock.h:302-309), gd_energy =
 * Energy (src/dock.h:108-128), gd_record = LigRecordSingleStep
 * (src/dock.h:361-367).
 */
#ifndef GEAUXDOCK_H
#define GEAUXDOCK_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Type-table dimensions (src/size.h:95-113). */
#define GD_MAXTP1 30 /* receptor point types */
#define GD_MAXTP2 24 /* ligand atom types */
#define GD_MAXTP3 50 /* electrostatic point classes */
#define GD_MAXTP4 20 /* residue types */
#define GD_MAXWEI 10 /* energy terms incl. total */
/* MCS field value marking "atom absent in this template" (src/size.h:142-143). */
#define GD_MCS_INVALID_COORD 728492.0f

/* Engine capacity limits (runtime checks; each violation returns GD_ERR_CAPACITY). */
#define GD_MAX_LNA 256     /* ligand atoms per complex */
#define GD_MAX_NTMP 64     /* temperatures per ladder */
#define GD_MAX_REFCONTACT 256 /* reference contacts per ligand atom (MCC map) */

typedef enum {
  GD_OK = 0,
  GD_ERR_INVALID = 1,  /* bad argument or inconsistent description */
  GD_ERR_CAPACITY = 2, /* a size exceeds an engine limit */
  GD_ERR_HIP = 3,      /* HIP runtime error (message in gd_last_error) */
  GD_ERR_STATE = 4,    /* call not valid in the current context state */
  GD_ERR_NOMEM = 5
} gd_status;

/* Energy parameters after OptimizeEnepara (src/util.C:717-768).
 * Two-index tables are [ligand type][receptor type]. */
typedef struct gd_enepara {
  float p1a[GD_MAXTP2][GD_MAXTP1];
  float p2a[GD_MAXTP2][GD_MAXTP1];
  float lj0, lj1;
  float el0;
  float el1;
  float ele[GD_MAXTP3];
  float a1; /* 4 - 3*el0 */
  float b1; /* 2*el0 - 3 */
  float pmf0[GD_MAXTP2][GD_MAXTP1];
  float pmf1[GD_MAXTP2][GD_MAXTP1];
  float hdb0[GD_MAXTP2][GD_MAXTP1];
  float hdb1[GD_MAXTP2][GD_MAXTP1];
  float hpp[GD_MAXTP4];
  float hpl0[GD_MAXTP2];
  float hpl1[GD_MAXTP2];
  float hpl2[GD_MAXTP2];
  float kde2; /* -0.5 / bw^2 */
  float kde3; /* (bw * sqrt(2 pi))^3 */
  float w[GD_MAXWEI];
  float a_para[GD_MAXWEI];
  float b_para[GD_MAXWEI];
} gd_enepara;

typedef struct gd_temp {
  float t;
  float minus_beta; /* -1 / (k_B T), k_B = 1 */
  int order;
} gd_temp;

typedef struct gd_replica {
  int idx_rep; /* flat index n_tmp*n_lig*prt + n_lig*tmp + lig */
  int idx_prt;
  int idx_tmp; /* current temperature index (changes on exchange) */
  int idx_lig;
} gd_replica;

/* e[0..8] hold the normalized, unweighted terms
 * (vdw, ele, pmf, psp, hdb, hpc, kde, lhm, dst); e[9] the weighted total. */
typedef struct gd_energy {
  float e[GD_MAXWEI];
  float cms;
  float rmsd;
} gd_energy;

typedef struct gd_record {
  gd_replica replica;
  gd_energy energy;
  float movematrix[6]; /* translation x y z, Euler angles x y z */
  int step;
} gd_record;

/* Move proposal kinds. The reference draws scale*(2U-1) per DOF
 * (src/kernel_cuda_l1_montecarlo.cu:100, src/kernel_cuda_l2_move.cu:27-30). */
#define GD_MOVE_UNIFORM 0
#define GD_MOVE_GAUSSIAN 1 /* scale * N(0,1) per DOF */

/* One complex: every ligand conformer x protein conformer x temperature.
 * All arrays are host pointers, copied by gd_add_complex.
 * Per-point receptor attributes are shared by all protein conformers; only
 * coordinates differ (conformers are models of one structure). */
typedef struct gd_complex_desc {
  int n_lig, n_prt, n_tmp; /* ComplexSize (src/dock.h:87-100) */
  int lna, pnp, pnk, pos;
  int n_res; /* PSP columns: residue sequence numbers seq3r lie in [0, n_res) */
  int n_mcs; /* fields per MCS template, indexed by ligand atom number + 1 */

  /* Ligand conformers in the ligand frame (Ligand.coord_orig, already centred). */
  const float *lig_x, *lig_y, *lig_z; /* [n_lig * lna] */
  const int *lig_t;                   /* [lna] atom type in [0, GD_MAXTP2) */
  const float *lig_c;                 /* [lna] charge */
  const int *lig_n;                   /* [lna] atom number; MCS index is n + 1 */
  float pocket_center[3];

  /* Protein conformers (Protein after OptimizeProtein, src/util.C:574-646). */
  const float *prt_x, *prt_y, *prt_z; /* [n_prt * pnp] */
  const int *prt_t;                   /* [pnp] point type in [0, GD_MAXTP1) */
  const int *prt_c;                   /* [pnp] point class */
  const float *prt_ele;               /* [pnp] electrostatic parameter */
  const int *prt_seq3r;               /* [pnp] residue sequence number */
  const int *prt_c0_and_d12_or_c2;    /* [pnp] hydrophobic-contact flag */
  const float *prt_hpp;               /* [pnp] residue hydrophobicity */

  const float *psp; /* [GD_MAXTP2 * n_res]: psp[lig_t * n_res + seq3r] */

  const float *kde_x, *kde_y, *kde_z; /* [pnk] */
  const int *kde_t;                   /* [pnk] */

  const float *mcs_x, *mcs_y, *mcs_z; /* [pos * n_mcs] */
  const float *mcs_tcc;               /* [pos] */

  const gd_enepara *enepara;
  const gd_temp *temp; /* [n_tmp] */

  float move_scale[6]; /* McPara.move_scale: translation x y z, rotation x y z */
  int move_kind;       /* GD_MOVE_UNIFORM or GD_MOVE_GAUSSIAN */
} gd_complex_desc;

/* Engine options. Zero-initialise, then set what you need. */
typedef struct gd_opts {
  int device;          /* HIP device ordinal */
  uint64_t seed;       /* Philox4x32-10 key; the whole run is a function of it */
  int steps_per_dump;  /* accepted-state records are drained every this many steps (0 -> 2000) */
  int record;          /* 1 = keep accepted-state records (src/kernel_cuda_l3_util.cu:73-108) */
  int lanes_per_replica; /* 0 = auto; else 8, 16, 32 or 64 */
  int atoms_per_lane;    /* 0 = auto; else 1..4 */
  int verbose;
  int receptor_split;    /* lane groups per replica, each over its share of every
                            receptor chunk and of every atom's KDE points and reference
                            contacts and of the MCS templates: 1, 2, 4 or 8 for every
                            workgroup; 0 = auto
                            per workgroup: whole ladders first, then the split that fills
                            the workgroup to 8 waves (ladders left over from an odd count
                            take 4 or 8 groups) */
  int ladders_per_pack;  /* 0 = auto; else temperature ladders per workgroup (replicas
                            per workgroup = this x n_tmp), capped by 512 lanes */
  int launch;            /* GD_LAUNCH_* flags; 0 = graph-captured, non-persistent */
  int receptor_tile;     /* 0 = each receptor conformer resident in LDS when it fits,
                            else streamed in tiles of 256 points; > 0 = every receptor
                            streamed through LDS in tiles of this many points */
} gd_opts;

/* gd_opts.launch flags. MC windows are captured once into a HIP graph (all
 * launch classes, forked over side streams) and replayed; the first step of
 * a window is read from device memory. PERSISTENT launches each class with
 * as many workgroups as stay resident, taking packs from a device counter in
 * longest-first order. SPREAD packing, for workloads with few temperature
 * ladders: every workgroup holds one wave of replicas (a ladder may span
 * several workgroups) at a receptor split of 8, and the replica exchanges run
 * in a kernel of their own between launches of steps_per_exchange steps, so
 * a small workload occupies more CUs. Energies then carry the bits of the
 * split of 8 (gd_calc_energy follows). It is automatic when the complexes
 * hold at most one ladder per two CUs and receptor_split and
 * ladders_per_pack are 0; SPREAD asks for it (not with a fixed
 * receptor_split or ladders_per_pack), NO_SPREAD keeps every ladder in one
 * workgroup. */
#define GD_LAUNCH_NO_GRAPH 1
#define GD_LAUNCH_PERSISTENT 2
#define GD_LAUNCH_SPREAD 4
#define GD_LAUNCH_NO_SPREAD 8

typedef struct gd_run_stats {
  double t_compute; /* seconds in MC kernels (HIP events), the reference's McLog.t0 */
  double t_wall;    /* seconds including record transfer, McLog.t1 */
  int64_t steps;    /* MC steps this call */
  int64_t replica_steps;
  int64_t accepted;          /* accepted MC moves this call */
  int64_t exchanges_accepted;
  int kernel_launches;
  double kernel_ms_avg;      /* mean MC-kernel duration (HIP events) */
} gd_run_stats;

/* Record sink: called once per dump window and complex with that window's
 * accepted-state records, ordered by replica then step. recs is valid until
 * the sink returns (the library reuses its pinned staging buffer). */
typedef void (*gd_record_sink)(void *user, int complex_id, const gd_record *recs, int64_t n);

typedef struct gd_ctx gd_ctx;

/* Create a context on opts->device. Replaces the set-up half of Run()
 * (src/run.cu:57-134). */
gd_status gd_create(gd_ctx **out, const gd_opts *opts);
void gd_destroy(gd_ctx *ctx);
const char *gd_last_error(const gd_ctx *ctx);

/* Copy one complex to the device (src/run.cu:138-223). Returns its id.
 * Replicas are laid out as SetReplica does (src/util.C:842-861). */
gd_status gd_add_complex(gd_ctx *ctx, const gd_complex_desc *desc, int *complex_id);

/* Run n_steps MC steps for every replica of every complex, with a replica-
 * exchange attempt every steps_per_exchange steps (0 = no exchange).
 * The first call first evaluates every replica's start pose and builds each
 * complex's reference contact map (MonteCarlo_Init_d,
 * src/kernel_cuda_l1_montecarlo.cu:12-68). Replaces the launch loop of
 * src/kernel_cuda_lancher_siglegpu.cu:5-69. sink may be NULL. */
gd_status gd_run(gd_ctx *ctx, int64_t n_steps, int steps_per_exchange,
                 gd_record_sink sink, void *user, gd_run_stats *stats);

/* Per-replica state of one complex (n_rep = n_lig*n_prt*n_tmp entries each;
 * any pointer may be NULL). acs_mc counts accepted MC moves, acs_x accepted
 * exchanges per replica (src/kernel_cuda_l1_exchangereplicas.cu:65-66 keeps
 * the same counts per temperature slot). */
gd_status gd_get_state(gd_ctx *ctx, int complex_id, float *movematrix /*[n_rep*6]*/,
                       gd_energy *energy_old, gd_replica *replica, int32_t *acs_mc,
                       int32_t *acs_x);

/* Checkpoint / restore of the whole MC state (poses, energies, temperature
 * assignment, counters, step counter). The RNG is counter based, so a resumed
 * run continues bit-identically. size: in = buffer size, out = bytes needed. */
gd_status gd_checkpoint(gd_ctx *ctx, void *buf, size_t *size);
gd_status gd_restore(gd_ctx *ctx, const void *buf, size_t size);

/* Parity hook: energies of n poses of one complex, through the same device
 * code as the MC kernel (CalcEnergy_d + CalcRmsd_d + CalcMcc_d,
 * src/kernel_cuda_l2_calcenergy.cu:19-360). movematrix is [n*6]. Needs the
 * complex's reference contact map (built here on first use). Each pose is
 * summed with the receptor split of its ligand conformer's MC workgroups, so
 * an MC step and this call give a pose the same bits (before the first
 * gd_run: the split the complexes added so far would get). */
gd_status gd_calc_energy(gd_ctx *ctx, int complex_id, int n, const int *idx_lig,
                         const int *idx_prt, const float *movematrix, gd_energy *out);

/* gd_calc_energy without any device allocation, for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc). No counterpart in the reference. Same inputs, outputs
 * and kernels as gd_calc_energy; the pose tables and energies of the call are
 * staged in one file-scope device array of GD_NOALLOC_SCRATCH_BYTES per GPU
 * (held until the process ends). At most GD_NOALLOC_MAX_POSES poses per call
 * (GD_ERR_CAPACITY beyond). The complex table and contact map must already be
 * on the device: one gd_calc_energy or gd_run call after the last
 * gd_add_complex (GD_ERR_STATE otherwise). Every call on a GPU shares that
 * GPU's array, whatever the context: calls that overlap on one GPU race. */
#define GD_NOALLOC_SCRATCH_BYTES (4 << 20)
#define GD_NOALLOC_MAX_POSES 32768
gd_status gd_calc_energy_noalloc(gd_ctx *ctx, int complex_id, int n, const int *idx_lig,
                                 const int *idx_prt, const float *movematrix, gd_energy *out);

/* Parity hook: Metropolis decisions accept[i] = u[i] < exp((e_new[i]-e_old[i]) *
 * minus_beta[i]) on the device (Accept_d, src/kernel_cuda_l2_accept.cu:16-42). */
gd_status gd_accept(gd_ctx *ctx, int n, const float *e_new, const float *e_old,
                    const float *minus_beta, const float *u, int32_t *accept);

/* Parity hook: one temperature-exchange sweep over n_groups ladders of n_tmp
 * slots on the device (ExchangeReplicas_d, src/kernel_cuda_l1_exchangereplicas.cu:29-78).
 * etot, idx_tmp, acs: [n_groups*n_tmp] by slot; minus_beta: [n_tmp] by
 * temperature; u: [n_groups*(n_tmp/2)] one uniform per attempted pair.
 * idx_tmp and acs are updated in place. */
gd_status gd_exchange(gd_ctx *ctx, int n_groups, int n_tmp, int mode_t, const float *etot,
                      const float *minus_beta, const float *u, int32_t *idx_tmp,
                      int32_t *acs);

/* Parity hook: the device's Philox4x32-10 stream. For i in [0,n):
 * out[4i..4i+3] = philox4x32_10(ctr[4i..4i+3], key[2i..2i+1]). */
gd_status gd_philox(gd_ctx *ctx, int n, const uint32_t *ctr, const uint32_t *key, uint32_t *out);

/* Number of replicas of a complex (n_lig*n_prt*n_tmp). */
int gd_num_replicas(const gd_ctx *ctx, int complex_id);
/* Lowest total energy each replica of a complex has held since the start
 * poses (or the last gd_reset_best), and that pose: best_e[n_rep],
 * best_movematrix[n_rep * 6] (either may be NULL). No reference counterpart:
 * the reference keeps every accepted state (src/kernel_cuda_l3_util.cu:73-108)
 * and selects poses from those records afterwards (src/kgs.C); the engine
 * also keeps the running minimum on the device so runs need not record. */
gd_status gd_get_best(gd_ctx *ctx, int complex_id, float *best_e, float *best_movematrix);
/* Best-so-far := each replica's current state. */
gd_status gd_reset_best(gd_ctx *ctx);
/* Lane shape chosen for a complex: lanes per replica and ligand atoms per
 * lane (the MC kernel instance that runs it). Under spread packing, unless
 * lanes_per_replica or atoms_per_lane is forced, the shape is widened once,
 * when the complex first goes to the device (first gd_calc_energy or
 * gd_run), to the one that makes the most useful one-wave packs; it is fixed
 * from then on. No reference counterpart: the reference fixes its launch
 * shape at compile time (src/Makefile:77-105). */
gd_status gd_complex_shape(const gd_ctx *ctx, int complex_id, int *lanes_per_replica,
                           int *atoms_per_lane);

/* ---- learned energy surrogate -------------------------------------------
 * No counterpart in the reference. A small MLP (56 -> 128 -> 128 -> 1, SiLU,
 * summed over ligand atoms) on per-atom features: one-hot ligand type (24) and
 * 32 radial counts of receptor points by class (t mod 4) at 1..8 A with a
 * smooth 8 A cutoff, read from per-protein-conformer grids. Trained on the GPU
 * with BF16 MFMA forward/backward kernels (FP32 accumulate) and FP32 Adam.
 * Parameter vector (FP32, gd_sur_num_params() floats): W1[64][128] (row 56 is
 * the first-layer bias, rows 57..63 unused), W2[128][128], b2[128], w3[128], b3. */
int gd_sur_num_params(void);
/* Feature grids around each complex's pocket (spacing, half width in A). */
gd_status gd_sur_build_grids(gd_ctx *ctx, float spacing, float half_width);
/* Features of n poses: out[(pose*lna + atom)*56 + f]. */
gd_status gd_sur_features(gd_ctx *ctx, int complex_id, int n, const int *idx_lig,
                          const int *idx_prt, const float *movematrix, float *out);
gd_status gd_sur_set_params(gd_ctx *ctx, const float *params);
gd_status gd_sur_get_params(gd_ctx *ctx, float *params);
/* Training set held on the device: poses of a complex with their targets. */
gd_status gd_sur_dataset_clear(gd_ctx *ctx);
gd_status gd_sur_dataset_reserve(gd_ctx *ctx, int64_t rows);
gd_status gd_sur_dataset_add(gd_ctx *ctx, int complex_id, int n, const int *idx_lig,
                             const int *idx_prt, const float *movematrix, const float *target);
/* The same for poses of several complexes, in the given order (complex_id[n]). */
gd_status gd_sur_dataset_add_multi(gd_ctx *ctx, int n, const int *complex_id, const int *idx_lig,
                                   const int *idx_prt, const float *movematrix,
                                   const float *target);
int64_t gd_sur_dataset_size(const gd_ctx *ctx);
/* Mean squared error over poses [pose_begin, pose_end) and its gradient. */
gd_status gd_sur_loss_grad(gd_ctx *ctx, int pose_begin, int pose_end, double *loss, float *grad);
gd_status gd_sur_predict(gd_ctx *ctx, int pose_begin, int pose_end, float *out);
/* Delayed-acceptance MC (surrogate transition, no counterpart in the
 * reference): each group of k_sub steps runs k_sub Metropolis steps on the
 * surrogate total S = sur_scale * (sum over ligand atoms of the MLP) / lna
 * + the exact KDE, MCS and pocket-distance terms, then accepts the composite
 * move with min(1, exp(-beta [(E(y) - E(x)) -
 * (S(y) - S(x))])) from one exact evaluation, so the chain keeps the exact
 * Boltzmann distribution. n_steps and steps_per_exchange are multiples of
 * k_sub; replica-steps count surrogate steps. For a complex with at least 256
 * KDE points or 32 MCS templates, the groups of a receptor split share the
 * surrogate's KDE and MCS sums (S then differs from the whole-list sum by
 * summation order only). */
gd_status gd_run_da(gd_ctx *ctx, int64_t n_steps, int steps_per_exchange, int k_sub,
                    float sur_scale, gd_record_sink sink, void *user, gd_run_stats *stats);
/* Tabulate the network on every complex's surrogate grid: for each ligand
 * atom type, the output at each grid point (features = the grid's values),
 * evaluated with the BF16 MFMA forward kernel. Until the parameters change,
 * gd_run_da and gd_sur_surrogate then read the surrogate by trilinear
 * interpolation of these tables (8 loads per atom) instead of evaluating
 * the network per atom. No reference counterpart. */
gd_status gd_sur_tabulate(gd_ctx *ctx);
/* Surrogate total S of n poses exactly as the delayed-acceptance kernel
 * computes it (tabulated or direct), out[n]. No reference counterpart. */
gd_status gd_sur_surrogate(gd_ctx *ctx, int complex_id, int n, const int *idx_lig,
                           const int *idx_prt, const float *movematrix, float sur_scale,
                           float *out);
/* Adam over contiguous batches of batch_poses poses. epoch_loss[epochs]
 * (optional): for each epoch the mean squared residual of its batches, each
 * taken before its batch's update; the last entry is the loss over the whole
 * range with the final parameters. */
gd_status gd_sur_train(gd_ctx *ctx, int pose_begin, int pose_end, int epochs, int batch_poses,
                       float lr, double *epoch_loss, double *seconds);

/* ---- post-MC clustering (device k-means) ---------------------------------
 * clusterByKmeans (src/util.C:1801-1880) over seq_kmeans (src/seq_kmeans.c)
 * on the nine energy terms e[0..8] of each object: centers start at the first
 * k objects of a seeded random permutation, Lloyd passes run while the share
 * of objects that change cluster exceeds `threshold` (the reference: 0.001,
 * at most 501 passes), and each non-empty cluster reports its member nearest
 * the final center (ties: lower index). A group with fewer than k objects
 * reports every object as its own medoid. Cluster sums are exact (fixed
 * point), so results are bit-reproducible. Limits: k <= 512, 2^19 objects
 * per group. */

/* groups: objects [group_off[g], group_off[g+1]) of features (row-major,
 * 9 floats per object). medoid, size: [n_groups * k], group-local index or
 * -1 for an empty cluster; passes (optional): [n_groups]; device_ms
 * (optional): device time of the passes and the medoid search. */
gd_status gd_kmeans_medoids(gd_ctx *ctx, int n_groups, const int64_t *group_off,
                            const float *features, int k, float threshold, int max_pass,
                            uint64_t seed, int *medoid, int *size, int *passes,
                            double *device_ms);

/* post_mc (src/post_mc.C:40-101): the records of each replica (ascending
 * replica id, in the order given) clustered into k_replica clusters, then
 * all their medoids, in replica and cluster order, into k_final. Writes up to
 * k_final medoids as record indices with their level-2 cluster sizes;
 * device_ms (optional): device time of both levels. */
gd_status gd_post_mc(gd_ctx *ctx, const gd_record *recs, int64_t n, int k_replica, int k_final,
                     uint64_t seed, int64_t *medoid_rec, int *medoid_size, int *n_medoids,
                     double *device_ms);

#ifdef __cplusplus
}
#endif

#endif /* GEAUXDOCK_H */
