// gd_engine.hip -- host side of the C ABI in include/geauxdock.h.
//
// Replaces Run() and its single-GPU launcher (src/run.cu:33-289,
// src/kernel_cuda_lancher_siglegpu.cu:5-69): uploads complexes once, keeps
// all replica state resident in HBM, and advances every replica of every
// complex with one fused kernel launch per record window.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "gd_sur.hip"
#include "gd_util.hip"

using namespace gd;

// Device scratch of gd_calc_energy_noalloc: the pose tables and the energies
// of one call (packs, slots, conformer indices, movematrices, outputs), so
// that entry point allocates nothing. One copy per GPU, held until the
// process ends; host code reaches it only through the address
// hipGetSymbolAddress returns for the current device (gd_ctx::scratch).
static __device__ __align__(256) unsigned char scratch[GD_NOALLOC_SCRATCH_BYTES];

namespace {

// One kernel launch of a pack set: packs [begin, end) share the atoms-per-lane
// kernel instance, the LDS layout and the block shape.
struct LaunchClass {
  int apl = 1, begin = 0, end = 0;
  int tile = 0;      // receptor tile length (0: receptor resident in LDS)
  int block = 64;    // lanes: each pack's ns groups of its `half` lanes
  int half_max = 64;  // largest pack half (the block of an unsplit launch)
  size_t lds = 0;    // dynamic LDS bytes
  KParams lay{};     // LDS offsets and exchange-buffer fields
};

struct HostComplex {
  gd_complex_desc d;  // sizes, scalars (pointers are not kept)
  int n_rep = 0, rep_base = 0;
  int lpr = 0, apl = 0, n_chunks = 0;
  bool tiled = false;  // receptor streamed through LDS tiles
  bool ref_built = false;
  bool shape_final = false;  // lane shape fixed (the complex's energies have bits)
  std::vector<int> lig_t;  // ligand atom types (host copy)
  std::vector<int> lig_ns;  // receptor split of each ligand conformer's MC packs
  DevComplex dev{};
  std::vector<void *> bufs;  // device allocations owned by this complex
  size_t lds_bytes = 0;
};

}  // namespace

struct gd_ctx {
  gd_opts opts{};
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  // side streams for concurrent launches of the per-APL kernel classes
  hipStream_t aux[3] = {nullptr, nullptr, nullptr};
  hipEvent_t ev_fork = nullptr, ev_join[3] = {nullptr, nullptr, nullptr};
  std::string err;
  int n_cu = 0;
  size_t max_lds = 65536;
  unsigned char *scratch = nullptr;  // this device's copy of the file-scope scratch array
  std::vector<HostComplex> cx;
  DevComplex *d_cx = nullptr;
  bool cx_dirty = true;
  int *d_overflow = nullptr;
  // replica state
  bool state_alloc = false, initialized = false;
  int n_rep = 0;
  int64_t step = 0;
  float *d_mm = nullptr;
  gd_energy *d_eold = nullptr;
  float *d_best_e = nullptr, *d_best_mm = nullptr;
  int *d_idx_tmp = nullptr, *d_acs_mc = nullptr, *d_acs_x = nullptr, *d_rep_flat = nullptr;
  // record rings: the MC kernels append accepted states to ring `ring`; with
  // a sink, windows alternate between two rings so one window's records are
  // drained while the next window runs (ring 1 is allocated on first use)
  gd_record *d_rec = nullptr;  // ring 0
  int *d_rec_cnt = nullptr;
  gd_record *d_ring[2] = {nullptr, nullptr};
  int *d_ring_cnt[2] = {nullptr, nullptr};
  int ring = 0;
  int rec_cap = 0;
  hipStream_t drain_stream = nullptr;
  // record drains: device offsets and compacted window, and the pinned host
  // staging buffer the sink reads, grown as needed and reused across windows
  int64_t *d_drain_off = nullptr;
  gd_record *d_drain = nullptr, *h_drain = nullptr;
  int64_t drain_cap = 0;
  // packs of the MC launch
  Pack *d_packs = nullptr;
  int *d_slots = nullptr;
  int n_packs = 0, ladders_per_pack = 0;
  // spread packing: ladders exchanged between launches, (complex, ladder,
  // replica of slot 0, replica stride)
  int4 *d_ladders = nullptr;
  int n_ladders = 0;
  std::vector<LaunchClass> mc_classes;  // kernel launches of the MC pack set
  // device control words: [0] first step of a graph replay, [1 + i] pack
  // counter of class i in persistent launches
  int *d_ctrl = nullptr;
  int h_step = 0;  // host source of the step upload (outlives the async copy)
  // captured MC launch per record ring (one graph per window shape)
  struct GraphSlot {
    hipGraphExec_t exec = nullptr;
    bool valid = false;
    int mode = -1, win = 0, nc = 0, da = 0, tab = 0, persistent = 0;
    float scale = 0.0f;
  } gslot[2];
  // learned surrogate
  struct Sur {
    std::vector<float *> grid;  // per complex [n_prt][n^3][32]
    std::vector<SurGrid> g;
    std::vector<float *> tab;   // per complex [n_prt][NLT][n^3] network outputs
    bool tab_valid = false;     // tab holds the current parameters' outputs
    float *p = nullptr, *m = nullptr, *v = nullptr, *grad = nullptr;  // FP32 master + Adam
    unsigned short *W1T = nullptr, *W2T = nullptr, *W2R = nullptr;    // bf16 layouts
    int64_t adam_t = 0;
    bool have_params = false;
    // dataset (rows padded, last 64 rows zero)
    unsigned short *X = nullptr;  // feature rows [cap + 64][64] bf16
    int64_t rows = 0, cap = 0;
    std::vector<int64_t> off{0};  // pose row offsets
    std::vector<float> target;
    std::vector<int> pose_cx;
    int *pose_of_row = nullptr;
    int64_t *d_off = nullptr;
    float *d_target = nullptr;
    bool dirty = true;
    // work buffers
    float *y = nullptr, *dS = nullptr, *sq = nullptr, *S = nullptr;
    float *p1 = nullptr, *p2 = nullptr, *pb2 = nullptr, *pw3 = nullptr;
    float *dyr = nullptr;  // per-row output gradient of the current batch
    int64_t work_rows = 0;
    int last_groups = 0;  // backward workgroups (partials) of the last batch
  } sur;
};

namespace {

gd_status fail(gd_ctx *c, gd_status s, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (c) c->err = buf;
  return s;
}

#define HIPCHK(expr)                                                                  \
  do {                                                                                \
    hipError_t e_ = (expr);                                                           \
    if (e_ != hipSuccess)                                                             \
      return fail(ctx, GD_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(e_));    \
  } while (0)

template <class T>
gd_status dev_upload(gd_ctx *ctx, HostComplex &h, T **dst, const T *src, size_t n) {
  void *p = nullptr;
  const size_t bytes = std::max<size_t>(n, 1) * sizeof(T);
  HIPCHK(hipMalloc(&p, bytes));
  h.bufs.push_back(p);
  if (n) HIPCHK(hipMemcpyAsync(p, src, n * sizeof(T), hipMemcpyHostToDevice, ctx->stream));
  *dst = static_cast<T *>(p);
  return GD_OK;
}

#define TRY(expr)                      \
  do {                                 \
    gd_status s_ = (expr);             \
    if (s_ != GD_OK) return s_;        \
  } while (0)

// Lanes per replica and atoms per lane: the fewest lane slots lpr*apl that
// hold lna atoms with apl <= 4, preferring 2..3 atoms per lane on ties (more
// independent work per lane), with a whole ladder within 512 lanes.
void choose_shape(int lna, int n_tmp, int force_lpr, int force_apl, int *lpr, int *apl) {
  int best_l = 64, best_a = (lna + 63) / 64, best_cost = 1 << 30;
  const int cands[4] = {8, 16, 32, 64};
  for (int ci = 0; ci < 4; ++ci) {
    const int l = cands[ci];
    if (force_lpr && l != force_lpr) continue;
    const int a = (lna + l - 1) / l;
    if (a > 4 || (force_apl && a > force_apl)) continue;
    const int aa = force_apl ? force_apl : a;
    if (l * n_tmp > 512 && !force_lpr) continue;
    const int cost = l * aa * 8 + std::abs(aa - 2);
    if (cost < best_cost) {
      best_cost = cost;
      best_l = l;
      best_a = aa;
    }
  }
  *lpr = best_l;
  *apl = best_a;
}

// Default tile length of receptors that do not fit in LDS.
constexpr int DEFAULT_TILE = 256;

// LDS bytes of a complex's pack with `slots` replicas: the resident layout
// (tile = 0) holds the receptor conformer and KDE points, the tiled layout two
// receptor tiles.
size_t lds_need(const HostComplex &h, int slots, int tile) {
  const gd_complex_desc &d = h.d;
  const size_t pts = tile > 0 ? (size_t)2 * tile : (size_t)d.pnp;
  const size_t kde = tile > 0 ? 0 : (size_t)d.pnk * 16;
  // the receptor split's two exchange buffers of a 512-lane workgroup (any
  // pack of the complex may be split)
  const size_t xp = (size_t)2 * (11 * 512 / 2) * 4;
  return pts * 24 + PAR_FLOATS * 4 + kde + NLT * 8 + (size_t)h.n_chunks * 16 +
         (size_t)slots * 16 + xp + 256;
}

// Tile length of a complex's packs, 0 when its receptor stays resident.
int complex_tile(const gd_ctx *ctx, const HostComplex &h);

}  // namespace

extern "C" {

const char *gd_last_error(const gd_ctx *c) { return c ? c->err.c_str() : "null context"; }

gd_status gd_create(gd_ctx **out, const gd_opts *opts) {
  if (!out) return GD_ERR_INVALID;
  *out = nullptr;
  gd_ctx *ctx = new (std::nothrow) gd_ctx();
  if (!ctx) return GD_ERR_NOMEM;
  if (opts) ctx->opts = *opts;
  if (ctx->opts.steps_per_dump <= 0) ctx->opts.steps_per_dump = 2000;
  ctx->device = ctx->opts.device;
  int ndev = 0;
  hipError_t e = hipGetDeviceCount(&ndev);
  if (e != hipSuccess || ndev <= 0) {
    *out = ctx;
    return fail(ctx, GD_ERR_HIP, "no HIP device available (%s)", hipGetErrorString(e));
  }
  *out = ctx;
  HIPCHK(hipSetDevice(ctx->device));
  hipDeviceProp_t prop;
  HIPCHK(hipGetDeviceProperties(&prop, ctx->device));
  ctx->n_cu = prop.multiProcessorCount;
  ctx->max_lds = prop.sharedMemPerBlock;
  if (prop.maxSharedMemoryPerMultiProcessor > 0)
    ctx->max_lds = std::max<size_t>(ctx->max_lds, std::min<size_t>(prop.maxSharedMemoryPerMultiProcessor, 160 * 1024));
  if (ctx->opts.verbose)
    fprintf(stderr, "[geauxdock] %s: %d CUs, LDS/block %zu, LDS/CU %zu, regs/block %d\n",
            prop.gcnArchName, prop.multiProcessorCount, (size_t)prop.sharedMemPerBlock,
            (size_t)prop.maxSharedMemoryPerMultiProcessor, prop.regsPerBlock);
  HIPCHK(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
  HIPCHK(hipStreamCreateWithFlags(&ctx->drain_stream, hipStreamNonBlocking));
  HIPCHK(hipEventCreate(&ctx->ev0));
  HIPCHK(hipEventCreate(&ctx->ev1));
  HIPCHK(hipEventCreateWithFlags(&ctx->ev_fork, hipEventDisableTiming));
  for (int i = 0; i < 3; ++i) {
    HIPCHK(hipStreamCreateWithFlags(&ctx->aux[i], hipStreamNonBlocking));
    HIPCHK(hipEventCreateWithFlags(&ctx->ev_join[i], hipEventDisableTiming));
  }
  HIPCHK(hipMalloc(&ctx->d_overflow, sizeof(int)));
  HIPCHK(hipGetSymbolAddress((void **)&ctx->scratch, HIP_SYMBOL(scratch)));
  return GD_OK;
}

void gd_destroy(gd_ctx *ctx) {
  if (!ctx) return;
  if (ctx->stream) {
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
  }
  for (auto &h : ctx->cx)
    for (void *p : h.bufs) (void)hipFree(p);
  void *ps[] = {ctx->d_cx, ctx->d_overflow, ctx->d_mm, ctx->d_eold, ctx->d_idx_tmp,
                ctx->d_acs_mc, ctx->d_acs_x, ctx->d_rep_flat, ctx->d_rec, ctx->d_rec_cnt,
                ctx->d_packs, ctx->d_slots, ctx->d_best_e, ctx->d_best_mm, ctx->d_ctrl,
                ctx->d_drain_off, ctx->d_drain, ctx->d_ring[1], ctx->d_ring_cnt[1],
                ctx->d_ladders};
  if (ctx->h_drain) (void)hipHostFree(ctx->h_drain);
  for (auto &g : ctx->gslot)
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
  for (void *p : ps)
    if (p) (void)hipFree(p);
  {
    auto &S = ctx->sur;
    for (float *g : S.grid)
      if (g) (void)hipFree(g);
    for (float *g : S.tab)
      if (g) (void)hipFree(g);
    void *qs[] = {S.p, S.m, S.v, S.grad, S.W1T, S.W2T, S.W2R, S.X, S.pose_of_row,
                  S.d_off, S.d_target, S.y, S.dS, S.sq, S.S, S.p1, S.p2, S.pb2, S.pw3,
                  S.dyr};
    for (void *p : qs)
      if (p) (void)hipFree(p);
  }
  if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
  if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
  if (ctx->ev_fork) (void)hipEventDestroy(ctx->ev_fork);
  for (int i = 0; i < 3; ++i) {
    if (ctx->ev_join[i]) (void)hipEventDestroy(ctx->ev_join[i]);
    if (ctx->aux[i]) (void)hipStreamDestroy(ctx->aux[i]);
  }
  if (ctx->drain_stream) (void)hipStreamDestroy(ctx->drain_stream);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

gd_status gd_complex_shape(const gd_ctx *ctx, int cid, int *lpr, int *apl) {
  if (!ctx || cid < 0 || cid >= (int)ctx->cx.size() || !lpr || !apl) return GD_ERR_INVALID;
  *lpr = ctx->cx[cid].lpr;
  *apl = ctx->cx[cid].apl;
  return GD_OK;
}

int gd_num_replicas(const gd_ctx *ctx, int cid) {
  if (!ctx || cid < 0 || cid >= (int)ctx->cx.size()) return -1;
  return ctx->cx[cid].n_rep;
}

gd_status gd_add_complex(gd_ctx *ctx, const gd_complex_desc *d, int *complex_id) {
  if (!ctx || !d) return GD_ERR_INVALID;
  if (ctx->state_alloc) return fail(ctx, GD_ERR_STATE, "complexes must be added before gd_run");
  if (d->n_lig <= 0 || d->n_prt <= 0 || d->n_tmp <= 0 || d->lna <= 0 || d->pnp <= 0 ||
      d->pnk < 0 || d->pos < 0)
    return fail(ctx, GD_ERR_INVALID, "non-positive size in complex description");
  if (d->lna > GD_MAX_LNA) return fail(ctx, GD_ERR_CAPACITY, "lna %d > %d", d->lna, GD_MAX_LNA);
  if (d->n_tmp > GD_MAX_NTMP)
    return fail(ctx, GD_ERR_CAPACITY, "n_tmp %d > %d", d->n_tmp, GD_MAX_NTMP);
  if (!d->lig_x || !d->lig_y || !d->lig_z || !d->lig_t || !d->lig_c || !d->lig_n || !d->prt_x ||
      !d->prt_y || !d->prt_z || !d->prt_t || !d->prt_c || !d->prt_ele || !d->prt_seq3r ||
      !d->prt_c0_and_d12_or_c2 || !d->prt_hpp || !d->enepara || !d->temp)
    return fail(ctx, GD_ERR_INVALID, "null array in complex description");
  if (d->pnk > 0 && (!d->kde_x || !d->kde_y || !d->kde_z || !d->kde_t))
    return fail(ctx, GD_ERR_INVALID, "null KDE array");
  if (d->pos > 0 && (!d->mcs_x || !d->mcs_y || !d->mcs_z || !d->mcs_tcc || d->n_mcs <= 0))
    return fail(ctx, GD_ERR_INVALID, "null MCS array");
  const int lna = d->lna, pnp = d->pnp, pnk = d->pnk, pos = d->pos;
  for (int l = 0; l < lna; ++l) {
    if (d->lig_t[l] < 0 || d->lig_t[l] >= NLT)
      return fail(ctx, GD_ERR_INVALID, "ligand atom %d type %d out of range", l, d->lig_t[l]);
    if (pos > 0 && (d->lig_n[l] + 1 < 0 || d->lig_n[l] + 1 >= d->n_mcs))
      return fail(ctx, GD_ERR_INVALID, "ligand atom %d number %d outside MCS fields", l, d->lig_n[l]);
  }
  bool any_psp = false;
  for (int p = 0; p < pnp; ++p) {
    if (d->prt_t[p] < 0 || d->prt_t[p] >= NPT)
      return fail(ctx, GD_ERR_INVALID, "receptor point %d type %d out of range", p, d->prt_t[p]);
    if (d->prt_c[p] == 2) {
      any_psp = true;
      if (d->prt_seq3r[p] < 0 || d->prt_seq3r[p] >= d->n_res)
        return fail(ctx, GD_ERR_INVALID, "receptor point %d seq3r %d outside [0,%d)", p,
                    d->prt_seq3r[p], d->n_res);
    }
  }
  if (any_psp && !d->psp) return fail(ctx, GD_ERR_INVALID, "null PSP table");
  for (int k = 0; k < pnk; ++k)
    if (d->kde_t[k] < 0 || d->kde_t[k] >= NLT)
      return fail(ctx, GD_ERR_INVALID, "KDE point %d type out of range", k);

  if (!(d->enepara->el1 > 0.0f) || !(d->enepara->lj0 > 0.0f))
    return fail(ctx, GD_ERR_INVALID, "el1 and lj0 must be positive");
  HIPCHK(hipSetDevice(ctx->device));
  ctx->cx.emplace_back();
  HostComplex &h = ctx->cx.back();
  const int cid = (int)ctx->cx.size() - 1;
  h.d = *d;
  h.n_rep = d->n_lig * d->n_prt * d->n_tmp;
  h.lig_t.assign(d->lig_t, d->lig_t + d->lna);
  choose_shape(lna, d->n_tmp, ctx->opts.lanes_per_replica, ctx->opts.atoms_per_lane, &h.lpr,
               &h.apl);
  if (h.lpr * h.apl < lna || h.apl < 1 || h.apl > 4)
    return fail(ctx, GD_ERR_INVALID, "no lane shape for lna %d (lpr %d apl %d)", lna, h.lpr, h.apl);

  // receptor sorted by (type, kind); kind: 2 = class 2, 1 = hydrophobic flag
  std::vector<int> kind(pnp), perm(pnp);
  for (int p = 0; p < pnp; ++p) {
    kind[p] = d->prt_c[p] == 2 ? KIND_PSP : (d->prt_c0_and_d12_or_c2[p] == 1 ? KIND_HPC : KIND_PLAIN);
    perm[p] = p;
  }
  std::stable_sort(perm.begin(), perm.end(), [&](int a, int b) {
    if (d->prt_t[a] != d->prt_t[b]) return d->prt_t[a] < d->prt_t[b];
    return kind[a] < kind[b];
  });
  std::vector<float4> xyze((size_t)d->n_prt * pnp);
  for (int c = 0; c < d->n_prt; ++c)
    for (int i = 0; i < pnp; ++i) {
      const int p = perm[i];
      const size_t o = (size_t)c * pnp;
      // the electrostatic parameter is stored divided by el1 (far-field form)
      xyze[o + i] = make_float4(d->prt_x[o + p], d->prt_y[o + p], d->prt_z[o + p],
                                d->prt_ele[p] / d->enepara->el1);
    }
  std::vector<float> hpp(pnp);
  std::vector<int> pspi(pnp);
  for (int i = 0; i < pnp; ++i) {
    const int p = perm[i];
    hpp[i] = d->prt_c0_and_d12_or_c2[p] == 1 ? d->prt_hpp[p] : 0.0f;
    pspi[i] = d->prt_c[p] == 2 ? d->prt_seq3r[p] : -1;
  }
  std::vector<int4> chunks;
  for (int i = 0; i < pnp;) {
    int j = i;
    const int t = d->prt_t[perm[i]], k = kind[perm[i]];
    while (j < pnp && d->prt_t[perm[j]] == t && kind[perm[j]] == k) ++j;
    chunks.push_back(make_int4(i, j - i, t, k));
    i = j;
  }
  h.n_chunks = (int)chunks.size();

  // type-pair parameter records [receptor type][ligand type]
  const gd_enepara &ep = *d->enepara;
  std::vector<float> par(PAR_FLOATS, 0.0f);
  const float hsc = sqrtf(0.5f * 1.44269504088896340736f);
  const float sqrt_2_pi_m1 = -1.0f / sqrtf(2.0f * 3.1415926535f);
  for (int pt = 0; pt < NPT; ++pt)
    for (int lt = 0; lt < NLT; ++lt) {
      float *r = &par[(pt * NLT + lt) * NPAR];
      r[P_P1A] = ep.p1a[lt][pt];
      r[P_P2A] = ep.p2a[lt][pt];
      r[P_PMF0] = ep.pmf0[lt][pt];
      r[P_PMF1] = ep.pmf1[lt][pt];
      const float h0 = ep.hdb0[lt][pt], h1 = ep.hdb1[lt][pt];
      const bool on = h0 > 0.1f;
      r[P_HA] = on ? h1 * hsc : 0.0f;
      r[P_HB] = on ? -h0 * h1 * hsc : 0.0f;
      r[P_HC] = on ? sqrt_2_pi_m1 * h1 : 0.0f;
      // the per-chunk constants of the pair loop, with its FP32 expressions
      const float p1a = r[P_P1A], pmf0 = r[P_PMF0];
      const float rn = fmaxf(pmf0, 1.0f / ep.el1) * 1.0001f;
      r[P_NR2] = rn * rn;
      r[P_A0] = p1a * ep.lj0;
      r[P_A1] = p1a * (ep.lj0 * ep.lj1);
      r[P_PB] = LOG2E * fmaf(0.5f, pmf0, 6.0f);
      r[P_PC] = -6.0f * LOG2E * pmf0;
    }

  // ligand conformers [n_lig][3][lna]
  std::vector<float> orig((size_t)d->n_lig * 3 * lna);
  for (int c = 0; c < d->n_lig; ++c)
    for (int l = 0; l < lna; ++l) {
      const size_t o = (size_t)c * lna + l;
      orig[(size_t)c * 3 * lna + l] = d->lig_x[o];
      orig[(size_t)c * 3 * lna + lna + l] = d->lig_y[o];
      orig[(size_t)c * 3 * lna + 2 * lna + l] = d->lig_z[o];
    }
  // MCS fields by atom: [pos][lna]
  std::vector<float4> mcs((size_t)std::max(pos, 1) * lna);
  std::vector<float> mcnt(std::max(pos, 1), 0.0f), tcc(std::max(pos, 1), 0.0f);
  for (int m = 0; m < pos; ++m) {
    tcc[m] = d->mcs_tcc[m];
    for (int l = 0; l < lna; ++l) {
      const size_t o = (size_t)m * d->n_mcs + d->lig_n[l] + 1;
      const bool ok = d->mcs_x[o] != GD_MCS_INVALID_COORD;
      mcs[(size_t)m * lna + l] = make_float4(d->mcs_x[o], d->mcs_y[o], d->mcs_z[o], ok ? 1.0f : 0.0f);
      mcnt[m] += ok ? 1.0f : 0.0f;
    }
  }
  // MCS moments per (ligand conformer, template) for the closed-form
  // restraint of the surrogate path (double sums, stored as float)
  std::vector<float4> mcsq((size_t)std::max(d->n_lig * pos, 1) * 4, make_float4(0, 0, 0, 0));
  for (int c = 0; c < d->n_lig; ++c)
    for (int m = 0; m < pos; ++m) {
      double so[3] = {0, 0, 0}, sm[3] = {0, 0, 0};
      int n = 0;
      for (int l = 0; l < lna; ++l) {
        const float4 f = mcs[(size_t)m * lna + l];
        if (f.w == 0.0f) continue;
        const size_t o = (size_t)c * 3 * lna + l;
        so[0] += orig[o];
        so[1] += orig[o + lna];
        so[2] += orig[o + 2 * lna];
        sm[0] += f.x;
        sm[1] += f.y;
        sm[2] += f.z;
        n += 1;
      }
      if (n == 0) continue;
      double ob[3], mb[3], A = 0.0, M[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
      for (int i = 0; i < 3; ++i) {
        ob[i] = so[i] / n;
        mb[i] = sm[i] / n;
      }
      for (int l = 0; l < lna; ++l) {
        const float4 f = mcs[(size_t)m * lna + l];
        if (f.w == 0.0f) continue;
        const size_t o = (size_t)c * 3 * lna + l;
        const double op[3] = {orig[o] - ob[0], orig[o + lna] - ob[1], orig[o + 2 * lna] - ob[2]};
        const double mp[3] = {f.x - mb[0], f.y - mb[1], f.z - mb[2]};
        for (int i = 0; i < 3; ++i) {
          A += op[i] * op[i] + mp[i] * mp[i];
          for (int j = 0; j < 3; ++j) M[i][j] += mp[i] * op[j];
        }
      }
      float4 *q = &mcsq[((size_t)c * pos + m) * 4];
      q[0] = make_float4((float)A, (float)ob[0], (float)ob[1], (float)ob[2]);
      q[1] = make_float4((float)mb[0], (float)mb[1], (float)mb[2], (float)M[0][0]);
      q[2] = make_float4((float)M[0][1], (float)M[0][2], (float)M[1][0], (float)M[1][1]);
      q[3] = make_float4((float)M[1][2], (float)M[2][0], (float)M[2][1], (float)M[2][2]);
    }
  // KDE sorted by type with per-type ranges
  std::vector<int> kperm(pnk);
  for (int k = 0; k < pnk; ++k) kperm[k] = k;
  std::stable_sort(kperm.begin(), kperm.end(), [&](int a, int b) { return d->kde_t[a] < d->kde_t[b]; });
  std::vector<float4> kde(std::max(pnk, 1));
  std::vector<int2> krange(NLT, make_int2(0, 0));
  for (int i = 0; i < pnk; ++i) {
    const int k = kperm[i];
    kde[i] = make_float4(d->kde_x[k], d->kde_y[k], d->kde_z[k], 0.0f);
  }
  for (int t = 0; t < NLT; ++t) {
    int b = 0;
    while (b < pnk && d->kde_t[kperm[b]] < t) ++b;
    int e = b;
    while (e < pnk && d->kde_t[kperm[e]] == t) ++e;
    krange[t] = make_int2(b, e);
  }

  DevComplex &D = h.dev;
  D.lna = lna;
  D.pnp = pnp;
  D.pnk = pnk;
  D.pos = pos;
  D.n_res = d->n_res;
  D.n_lig = d->n_lig;
  D.n_prt = d->n_prt;
  D.n_tmp = d->n_tmp;
  D.n_chunks = h.n_chunks;
  D.lpr = h.lpr;
  D.apl = h.apl;
  D.move_kind = d->move_kind;
  D.ref_total = 0;
  D.cid = cid;
  // a surrogate step evaluates the KDE and MCS terms exactly; for large sets
  // the receptor split's groups share them (one exchange per step)
  D.da_share = (d->pnk >= 256 || d->pos >= 32) ? 1 : 0;
  for (int i = 0; i < 3; ++i) D.pocket[i] = d->pocket_center[i];
  for (int i = 0; i < 6; ++i) D.move_scale[i] = d->move_scale[i];
  D.el0 = ep.el0;
  D.el1 = ep.el1;
  D.a1 = ep.a1;
  D.b1 = ep.b1;
  D.kde2l = ep.kde2 * 1.44269504088896340736f;
  D.kde3 = ep.kde3;
  for (int i = 0; i < GD_MAXWEI; ++i) {
    D.w[i] = ep.w[i];
    D.a_para[i] = ep.a_para[i];
    D.b_para[i] = ep.b_para[i];
  }
  for (int i = 0; i < NLT; ++i) {
    D.hpl0[i] = ep.hpl0[i];
    D.hpl1[i] = ep.hpl1[i];
    D.hpl2[i] = ep.hpl2[i];
  }
  for (int t = 0; t < GD_MAX_NTMP; ++t) D.minus_beta[t] = t < d->n_tmp ? d->temp[t].minus_beta : 0.0f;

  const float *psp_src = d->psp;
  std::vector<float> psp_zero;
  size_t psp_n = (size_t)NLT * std::max(d->n_res, 1);
  if (!psp_src) {
    psp_zero.assign(psp_n, 0.0f);
    psp_src = psp_zero.data();
  }
  TRY(dev_upload(ctx, h, (float4 **)&D.prt_xyze, xyze.data(), xyze.size()));
  TRY(dev_upload(ctx, h, (float **)&D.prt_hpp, hpp.data(), hpp.size()));
  TRY(dev_upload(ctx, h, (int **)&D.prt_psp, pspi.data(), pspi.size()));
  {
    std::vector<int> cls(pnp);
    for (int i = 0; i < pnp; ++i) cls[i] = d->prt_t[perm[i]] % 4;
    TRY(dev_upload(ctx, h, (int **)&D.prt_cls, cls.data(), cls.size()));
  }
  TRY(dev_upload(ctx, h, (int4 **)&D.chunks, chunks.data(), chunks.size()));
  TRY(dev_upload(ctx, h, (float **)&D.params, par.data(), par.size()));
  TRY(dev_upload(ctx, h, (float **)&D.lig_orig, orig.data(), orig.size()));
  TRY(dev_upload(ctx, h, (int **)&D.lig_t, d->lig_t, (size_t)lna));
  TRY(dev_upload(ctx, h, (float **)&D.lig_c, d->lig_c, (size_t)lna));
  TRY(dev_upload(ctx, h, (float4 **)&D.mcs, mcs.data(), mcs.size()));
  TRY(dev_upload(ctx, h, (float **)&D.mcs_tcc, tcc.data(), tcc.size()));
  TRY(dev_upload(ctx, h, (float **)&D.mcs_cnt, mcnt.data(), mcnt.size()));
  TRY(dev_upload(ctx, h, (float4 **)&D.mcsq, mcsq.data(), mcsq.size()));
  TRY(dev_upload(ctx, h, (float4 **)&D.kde, kde.data(), kde.size()));
  TRY(dev_upload(ctx, h, (int2 **)&D.kde_range, krange.data(), krange.size()));
  TRY(dev_upload(ctx, h, (float **)&D.psp, psp_src, psp_n));
  TRY(dev_upload<int2>(ctx, h, &D.ref, nullptr, 0));
  {
    void *p = nullptr;
    HIPCHK(hipMalloc(&p, (size_t)lna * KMAX * sizeof(int2)));
    h.bufs.push_back(p);
    D.ref = (int2 *)p;
    HIPCHK(hipMalloc(&p, (size_t)lna * sizeof(int)));
    h.bufs.push_back(p);
    D.ref_cnt = (int *)p;
  }
  HIPCHK(hipStreamSynchronize(ctx->stream));
  h.lds_bytes = lds_need(h, 512 / h.lpr, 0);
  h.tiled = ctx->opts.receptor_tile > 0 || h.lds_bytes > ctx->max_lds;
  if (h.tiled) {
    h.lds_bytes = lds_need(h, 512 / h.lpr, complex_tile(ctx, h));
    if (h.lds_bytes > ctx->max_lds)
      return fail(ctx, GD_ERR_CAPACITY, "complex needs %zu bytes of LDS with receptor tiles, "
                  "device has %zu", h.lds_bytes, ctx->max_lds);
  }
  ctx->cx_dirty = true;
  if (complex_id) *complex_id = cid;
  return GD_OK;
}

}  // extern "C"

namespace {

bool spread_packing(const gd_ctx *ctx);

// Packs of one wave of replicas (spread packing) for a complex at lpr lanes
// per replica.
int64_t spread_pack_count(const HostComplex &h, int lpr) {
  const int r = std::max(1, 64 / lpr);
  return (int64_t)h.d.n_prt * h.d.n_lig * ((h.d.n_tmp + r - 1) / r);
}

// Lane shapes under spread packing, fixed when a complex is first put on the
// device. Spread packs hold one wave of replicas, so more lanes per replica
// make more packs: a complex takes the shape with the fewest rounds of packs
// over the CUs times atoms per lane (the fewest packs on a tie), among the
// shapes whose ladder fits a workgroup. Forced shapes are kept.
gd_status widen_shapes(gd_ctx *ctx) {
  const bool widen = spread_packing(ctx) && ctx->opts.lanes_per_replica == 0 &&
                     ctx->opts.atoms_per_lane == 0;
  int64_t packs = 0;
  for (const HostComplex &h : ctx->cx)
    if (h.shape_final) packs += spread_pack_count(h, h.lpr);
  for (HostComplex &h : ctx->cx) {
    if (h.shape_final) continue;
    h.shape_final = true;
    if (!widen) continue;
    const int lna = h.d.lna;
    auto cost = [&](int l, int a) {
      return (double)((packs + spread_pack_count(h, l) + ctx->n_cu - 1) / ctx->n_cu) * a;
    };
    int best_l = h.lpr, best_a = h.apl;
    double best = cost(best_l, best_a);
    for (int l = 8; l <= 64; l *= 2) {
      const int a = (lna + l - 1) / l;
      if (a > 4 || l * h.d.n_tmp > 512) continue;
      const double c = cost(l, a);
      if (c < best || (c == best && l < best_l)) {
        best = c;
        best_l = l;
        best_a = a;
      }
    }
    if (best_l != h.lpr) {
      const int ol = h.lpr, oa = h.apl;
      h.lpr = best_l;
      h.apl = best_a;
      const size_t need = lds_need(h, 512 / h.lpr, h.tiled ? complex_tile(ctx, h) : 0);
      if (need > ctx->max_lds) {  // keep the shape add_complex chose
        h.lpr = ol;
        h.apl = oa;
      } else {
        h.lds_bytes = need;
        h.dev.lpr = h.lpr;
        h.dev.apl = h.apl;
      }
    }
    packs += spread_pack_count(h, h.lpr);
  }
  return GD_OK;
}

gd_status sync_complexes(gd_ctx *ctx) {
  if (!ctx->cx_dirty) return GD_OK;
  TRY(widen_shapes(ctx));
  if (ctx->d_cx) HIPCHK(hipFree(ctx->d_cx));
  ctx->d_cx = nullptr;
  for (auto &g : ctx->gslot) g.valid = false;  // captured launches hold the old pointer
  std::vector<DevComplex> v;
  for (auto &h : ctx->cx) v.push_back(h.dev);
  HIPCHK(hipMalloc(&ctx->d_cx, std::max<size_t>(v.size(), 1) * sizeof(DevComplex)));
  HIPCHK(hipMemcpyAsync(ctx->d_cx, v.data(), v.size() * sizeof(DevComplex), hipMemcpyHostToDevice,
                        ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  ctx->cx_dirty = false;
  return GD_OK;
}

// Contact map of replica 0's start pose for each complex that lacks one.
gd_status ensure_ref(gd_ctx *ctx) {
  TRY(sync_complexes(ctx));
  bool any = false;
  for (int c = 0; c < (int)ctx->cx.size(); ++c) {
    HostComplex &h = ctx->cx[c];
    if (h.ref_built) continue;
    HIPCHK(hipMemsetAsync(ctx->d_overflow, 0, sizeof(int), ctx->stream));
    hipLaunchKernelGGL(gd_ref_kernel, dim3((h.d.lna + 63) / 64), dim3(64), 0, ctx->stream,
                       ctx->d_cx, c, ctx->d_overflow);
    HIPCHK(hipGetLastError());
    int ovf = 0;
    std::vector<int> cnt(h.d.lna);
    HIPCHK(hipMemcpyAsync(&ovf, ctx->d_overflow, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipMemcpyAsync(cnt.data(), h.dev.ref_cnt, cnt.size() * sizeof(int), hipMemcpyDeviceToHost,
                          ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    if (ovf) return fail(ctx, GD_ERR_CAPACITY, "complex %d: more than %d reference contacts for one ligand atom", c, KMAX);
    int tot = 0;
    for (int v : cnt) tot += v;
    h.dev.ref_total = tot;
    h.ref_built = true;
    any = true;
  }
  if (any) {
    ctx->cx_dirty = true;
    TRY(sync_complexes(ctx));
  }
  return GD_OK;
}

int complex_tile(const gd_ctx *ctx, const HostComplex &h) {
  if (!h.tiled) return 0;
  return ctx->opts.receptor_tile > 0 ? ctx->opts.receptor_tile : DEFAULT_TILE;
}

// LDS layout for a launch covering packs with these maxima, in blocks of
// `block` lanes; `split` adds the partial-sum exchange buffers of packs with
// more than one lane group ([ns*half][4] HPC sums and [ns][half/8][12] replica
// sums, ns*half = block). With tile > 0 the receptor arrays are two tiles of
// `tile` points and KDE stays in HBM.
void layout(KParams &P, size_t *bytes, int pnp_max, int pnk_max, int nch_max, int slots_max,
            int block, bool split, int tile = 0) {
  auto al = [](size_t v) { return (v + 15) & ~(size_t)15; };
  P.tile = tile;
  if (tile > 0) {
    pnp_max = 2 * tile;
    pnk_max = 0;
  }
  size_t o = al((size_t)pnp_max * 16);
  P.off_hpp = (int)o;
  o = al(o + (size_t)pnp_max * 4);
  P.off_psp = (int)o;
  o = al(o + (size_t)pnp_max * 4);
  P.off_par = (int)o;
  o = al(o + (size_t)PAR_FLOATS * 4);
  P.off_kde = (int)o;
  o = al(o + (size_t)std::max(pnk_max, 1) * 16);
  P.off_kr = (int)o;
  o = al(o + NLT * 8);
  P.off_ch = (int)o;
  o = al(o + (size_t)std::max(nch_max, 1) * 16);
  P.off_xe = (int)o;
  o = al(o + (size_t)slots_max * 4);
  P.off_xi = (int)o;
  o = al(o + (size_t)slots_max * 4);
  P.off_xa = (int)o;
  o = al(o + (size_t)slots_max * 4);
  P.off_xo = (int)o;
  o = al(o + (size_t)slots_max * 4);
  // split exchange: 4 floats of per-atom HPC sums per lane, 12 floats of
  // partial sums per replica (at most block/8 replicas): 5.5 floats per lane
  P.xp_floats = split ? (11 * block) / 2 : 0;
  P.off_xp = (int)o;
  o = al(o + (size_t)2 * P.xp_floats * 4);
  P.off_q = (int)o;
  o = al(o + 16);
  *bytes = o;
}

struct PackSet {
  std::vector<Pack> packs;
  std::vector<int> slots;
  std::vector<LaunchClass> classes;
};

// Copy the LDS layout and split fields of a launch class into P.
void apply_layout(KParams &P, const KParams &L) {
  P.off_hpp = L.off_hpp;
  P.off_psp = L.off_psp;
  P.off_par = L.off_par;
  P.off_kde = L.off_kde;
  P.off_kr = L.off_kr;
  P.off_ch = L.off_ch;
  P.off_xe = L.off_xe;
  P.off_xi = L.off_xi;
  P.off_xa = L.off_xa;
  P.off_xo = L.off_xo;
  P.off_xp = L.off_xp;
  P.xp_floats = L.xp_floats;
  P.off_q = L.off_q;
  P.tile = L.tile;
}

// Group packs into launch classes by (atoms per lane, tile, block shape,
// workgroups per CU that the LDS allows), each class longest packs first so
// the hardware dispatcher (or the persistent queue) fills the CUs in LPT
// order. Every pack of a class has the class's block (its ns groups of `half`
// lanes), so no wave of a workgroup is idle. A class's LDS is the largest need
// among its packs: a class closes when the next pack would lower the
// workgroups per CU it allows. Workgroups of 512 lanes take a CU's register
// file alone (two waves per SIMD), so LDS does not split their classes.
void pack_finish(const gd_ctx *ctx, PackSet &ps) {
  const size_t n_pk = ps.packs.size();
  std::vector<int> res(n_pk), block(n_pk), apl(n_pk), tile(n_pk);
  std::vector<double> cost(n_pk);
  for (size_t i = 0; i < n_pk; ++i) {
    Pack &p = ps.packs[i];
    const HostComplex &h = ctx->cx[p.cx];
    p.half = ((p.n_slots * h.lpr + 63) / 64) * 64;
    block[i] = p.half * p.ns;
    apl[i] = h.apl;
    // the tile length depends on the options only (not on the pack's block),
    // so every launch of a complex sums its chunks in the same order
    tile[i] = complex_tile(ctx, h);
    cost[i] = (double)h.d.pnp * (double)p.half * h.apl;
    KParams L{};
    size_t need = 0;
    layout(L, &need, (h.d.pnp + 3) & ~3, h.d.pnk, h.n_chunks, p.n_slots, block[i], p.ns > 1,
           tile[i]);
    res[i] = block[i] >= 512
                 ? 1
                 : (int)std::max<size_t>(1, std::min<size_t>(512 / block[i],
                                                             ctx->max_lds / std::max<size_t>(need, 1)));
  }
  // all per-pack keys are indexed by the pack's position before sorting
  std::vector<int> order(n_pk);
  for (size_t i = 0; i < n_pk; ++i) order[i] = (int)i;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
    if (apl[a] != apl[b]) return apl[a] < apl[b];
    if (tile[a] != tile[b]) return tile[a] < tile[b];
    if (block[a] != block[b]) return block[a] > block[b];
    if (res[a] != res[b]) return res[a] < res[b];
    return cost[a] > cost[b];
  });
  std::vector<Pack> sorted;
  for (int i : order) sorted.push_back(ps.packs[i]);
  ps.packs.swap(sorted);
  ps.classes.clear();
  for (size_t b = 0; b < n_pk;) {
    const int o0 = order[b];
    LaunchClass lc;
    lc.apl = apl[o0];
    lc.begin = (int)b;
    lc.block = block[o0];
    lc.tile = tile[o0];
    int pnp_max = 0, pnk_max = 0, nch_max = 0, slots_max = 1, half_max = 64;
    bool split = false;
    size_t e = b;
    for (; e < n_pk; ++e) {
      const int oe = order[e];
      if (apl[oe] != lc.apl || tile[oe] != lc.tile || block[oe] != lc.block || res[oe] != res[o0])
        break;
      const Pack &p = ps.packs[e];
      const HostComplex &h = ctx->cx[p.cx];
      const int np = std::max(pnp_max, h.d.pnp), nk = std::max(pnk_max, h.d.pnk);
      const int nc = std::max(nch_max, h.n_chunks), ns = std::max(slots_max, p.n_slots);
      const bool sp = split || p.ns > 1;
      KParams L{};
      size_t need = 0;
      layout(L, &need, (np + 3) & ~3, nk, nc, ns, lc.block, sp, lc.tile);
      const size_t fit = ctx->max_lds / std::max<size_t>(need, 1);
      if (e > b && (fit < 1 || (res[o0] > 1 && fit < (size_t)res[o0]))) break;
      pnp_max = np;
      pnk_max = nk;
      nch_max = nc;
      slots_max = ns;
      split = sp;
      half_max = std::max(half_max, p.half);
    }
    lc.end = (int)e;
    lc.half_max = half_max;
    layout(lc.lay, &lc.lds, (pnp_max + 3) & ~3, pnk_max, nch_max, slots_max, lc.block, split,
           lc.tile);
    ps.classes.push_back(lc);
    b = e;
  }
}

// Ladders of one (complex, protein conformer) grouped into packs, in
// ligand-conformer order: {first ligand conformer, ladders, receptor split}.
// The same grouping serves every protein conformer of the complex.
//
// A fixed split (receptor_split 1, 2, 4 or 8) takes `ladders_per_pack`
// ladders per pack (default: about 256 lanes), within 512 lanes per workgroup.
// The automatic split (0) fills every workgroup to 8 waves, the two waves per
// SIMD that the kernels' registers allow: ladders of w waves are grouped by
// powers of two, at most `gmax` ladders (and 8 waves) per pack, and each pack
// takes the smallest split that fills 8 waves with its groups, so a ladder
// left over from an odd count runs with more groups instead of leaving waves
// idle. A ladder of 5 or more waves runs unsplit.
struct LadderGroup {
  int l0, ng, ns;
};

std::vector<LadderGroup> ladder_groups(const gd_ctx *ctx, const HostComplex &h, int gmax) {
  std::vector<LadderGroup> out;
  const int n_lig = h.d.n_lig, glanes = h.d.n_tmp * h.lpr;
  const int fixed = ctx->opts.receptor_split, ladders = ctx->opts.ladders_per_pack;
  if (fixed > 0) {
    int G = ladders > 0 ? ladders : std::max(1, 256 / glanes);
    while (G > 1 && G * glanes > 512 / fixed) --G;
    for (int l0 = 0; l0 < n_lig; l0 += G) out.push_back({l0, std::min(G, n_lig - l0), fixed});
    return out;
  }
  const int w = (glanes + 63) / 64;
  if (w >= 5) {
    for (int l0 = 0; l0 < n_lig; ++l0) out.push_back({l0, 1, 1});
    return out;
  }
  const int gcap = std::max(1, std::min(gmax, 8 / w));
  for (int l0 = 0; l0 < n_lig;) {
    int G = 1;
    while (2 * G <= std::min(gcap, n_lig - l0)) G *= 2;
    int ns = 1;
    while (2 * ns * G * w <= 8) ns *= 2;
    out.push_back({l0, G, ns});
    l0 += G;
  }
  return out;
}

// Ladders per pack of the automatic split for the context's complexes:
// `ladders_per_pack` when set, else the most (up to 8) that still leaves at
// least four packs per CU for the longest-first balancing, down to one ladder
// per pack (each then filled by its split) when the batch is small.
int auto_gmax(const gd_ctx *ctx) {
  if (ctx->opts.ladders_per_pack > 0) return ctx->opts.ladders_per_pack;
  int g = 8;
  for (; g > 1; g /= 2) {
    int64_t packs = 0;
    for (const HostComplex &h : ctx->cx) packs += (int64_t)h.d.n_prt * ladder_groups(ctx, h, g).size();
    if (packs >= (int64_t)4 * ctx->n_cu) break;
  }
  return g;
}

// Spread packing: on request, or by default when the workload has at most
// one temperature ladder per two CUs (ladder-per-workgroup packing would
// leave most CUs idle) and no fixed split or ladders per pack is set.
bool spread_packing(const gd_ctx *ctx) {
  const int L = ctx->opts.launch;
  if (L & GD_LAUNCH_SPREAD) return true;
  if ((L & GD_LAUNCH_NO_SPREAD) || ctx->opts.receptor_split > 0 || ctx->opts.ladders_per_pack > 0)
    return false;
  int64_t ladders = 0;
  for (const HostComplex &h : ctx->cx) ladders += (int64_t)h.d.n_prt * h.d.n_lig;
  return ladders > 0 && 2 * ladders <= ctx->n_cu;
}

// Receptor split of each ligand conformer's MC packs. The split sets the
// summation order, so a pose's energy bits follow it; gd_calc_energy uses it.
std::vector<int> lig_splits(const gd_ctx *ctx, const HostComplex &h) {
  std::vector<int> ns(h.d.n_lig, spread_packing(ctx) ? 8 : 1);
  if (spread_packing(ctx)) return ns;
  for (const LadderGroup &lg : ladder_groups(ctx, h, auto_gmax(ctx)))
    for (int g = 0; g < lg.ng; ++g) ns[lg.l0 + g] = lg.ns;
  return ns;
}

// Spread packing: each pack holds one wave of replicas of one ladder
// (64 / lanes-per-replica temperature slots, in slot order) at a receptor
// split of 8, so each workgroup is 8 waves; a ladder of more slots spans
// several packs, and its exchanges run between launches.
PackSet build_spread_packs(gd_ctx *ctx) {
  PackSet ps;
  for (int c = 0; c < (int)ctx->cx.size(); ++c) {
    HostComplex &h = ctx->cx[c];
    const int n_tmp = h.d.n_tmp, n_lig = h.d.n_lig;
    const int r = std::max(1, 64 / h.lpr);
    h.lig_ns.assign(n_lig, 8);
    for (int p = 0; p < h.d.n_prt; ++p)
      for (int l = 0; l < n_lig; ++l)
        for (int t0 = 0; t0 < n_tmp; t0 += r) {
          Pack pk{};
          pk.cx = c;
          pk.prt = p;
          pk.n_groups = 0;
          pk.n_slots = std::min(r, n_tmp - t0);
          pk.ns = 8;
          pk.slot_off = (int)ps.slots.size();
          for (int t = t0; t < t0 + pk.n_slots; ++t)
            ps.slots.push_back(h.rep_base + n_tmp * n_lig * p + n_lig * t + l);
          ps.packs.push_back(pk);
        }
  }
  pack_finish(ctx, ps);
  return ps;
}

// MC packs: the ladder groups of every (complex, protein conformer).
PackSet build_mc_packs(gd_ctx *ctx) {
  if (spread_packing(ctx)) return build_spread_packs(ctx);
  PackSet ps;
  const int gmax = auto_gmax(ctx);
  for (int c = 0; c < (int)ctx->cx.size(); ++c) {
    HostComplex &h = ctx->cx[c];
    const int n_tmp = h.d.n_tmp, n_lig = h.d.n_lig;
    const std::vector<LadderGroup> groups = ladder_groups(ctx, h, gmax);
    h.lig_ns.assign(n_lig, 1);
    for (const LadderGroup &lg : groups)
      for (int g = 0; g < lg.ng; ++g) h.lig_ns[lg.l0 + g] = lg.ns;
    for (int p = 0; p < h.d.n_prt; ++p)
      for (const LadderGroup &lg : groups) {
        Pack pk{};
        pk.cx = c;
        pk.prt = p;
        pk.n_groups = lg.ng;
        pk.n_slots = lg.ng * n_tmp;
        pk.ns = lg.ns;
        pk.slot_off = (int)ps.slots.size();
        for (int g = 0; g < lg.ng; ++g)
          for (int t = 0; t < n_tmp; ++t)
            ps.slots.push_back(h.rep_base + n_tmp * n_lig * p + n_lig * t + (lg.l0 + g));
        ps.packs.push_back(pk);
      }
  }
  pack_finish(ctx, ps);
  return ps;
}

gd_status alloc_state(gd_ctx *ctx) {
  if (ctx->state_alloc) return GD_OK;
  TRY(ensure_ref(ctx));
  int n = 0;
  for (auto &h : ctx->cx) {
    h.rep_base = n;
    n += h.n_rep;
  }
  if (n == 0) return fail(ctx, GD_ERR_STATE, "no complexes added");
  ctx->n_rep = n;
  HIPCHK(hipMalloc(&ctx->d_mm, (size_t)n * 6 * sizeof(float)));
  HIPCHK(hipMalloc(&ctx->d_eold, (size_t)n * sizeof(gd_energy)));
  HIPCHK(hipMalloc(&ctx->d_best_e, (size_t)n * sizeof(float)));
  HIPCHK(hipMalloc(&ctx->d_best_mm, (size_t)n * 6 * sizeof(float)));
  HIPCHK(hipMalloc(&ctx->d_idx_tmp, (size_t)n * sizeof(int)));
  HIPCHK(hipMalloc(&ctx->d_acs_mc, (size_t)n * sizeof(int)));
  HIPCHK(hipMalloc(&ctx->d_acs_x, (size_t)n * sizeof(int)));
  HIPCHK(hipMalloc(&ctx->d_rep_flat, (size_t)n * sizeof(int)));
  HIPCHK(hipMalloc(&ctx->d_rec_cnt, (size_t)n * sizeof(int)));
  ctx->d_ring_cnt[0] = ctx->d_rec_cnt;
  std::vector<int> flat(n), itmp(n);
  for (auto &h : ctx->cx)
    for (int f = 0; f < h.n_rep; ++f) {
      flat[h.rep_base + f] = f;
      itmp[h.rep_base + f] = (f / h.d.n_lig) % h.d.n_tmp;  // SetReplica, src/util.C:846-852
    }
  HIPCHK(hipMemcpyAsync(ctx->d_rep_flat, flat.data(), n * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(hipMemcpyAsync(ctx->d_idx_tmp, itmp.data(), n * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(hipMemsetAsync(ctx->d_mm, 0, (size_t)n * 6 * sizeof(float), ctx->stream));
  HIPCHK(hipMemsetAsync(ctx->d_eold, 0, (size_t)n * sizeof(gd_energy), ctx->stream));
  HIPCHK(hipMemsetAsync(ctx->d_acs_mc, 0, (size_t)n * sizeof(int), ctx->stream));
  HIPCHK(hipMemsetAsync(ctx->d_acs_x, 0, (size_t)n * sizeof(int), ctx->stream));
  HIPCHK(hipMemsetAsync(ctx->d_rec_cnt, 0, (size_t)n * sizeof(int), ctx->stream));
  if (ctx->opts.record) {
    ctx->rec_cap = ctx->opts.steps_per_dump + 1;
    HIPCHK(hipMalloc(&ctx->d_rec, (size_t)n * ctx->rec_cap * sizeof(gd_record)));
    ctx->d_ring[0] = ctx->d_rec;
  }
  // receptor split: automatic (0, per pack) or fixed
  const int ns = ctx->opts.receptor_split;
  if (ns != 0 && ns != 1 && ns != 2 && ns != 4 && ns != 8)
    return fail(ctx, GD_ERR_INVALID, "receptor_split must be 0, 1, 2, 4 or 8");
  for (auto &h : ctx->cx)
    if (ns > 1 && h.d.n_tmp * h.lpr > 512 / ns)
      return fail(ctx, GD_ERR_CAPACITY,
                  "receptor split %d needs ladders of at most %d lanes (a ladder has %d)", ns,
                  512 / ns, h.d.n_tmp * h.lpr);
  if ((ctx->opts.launch & GD_LAUNCH_SPREAD) && (ns > 0 || ctx->opts.ladders_per_pack > 0))
    return fail(ctx, GD_ERR_INVALID,
                "GD_LAUNCH_SPREAD sets its own packs: leave receptor_split and ladders_per_pack 0");
  PackSet ps = build_mc_packs(ctx);
  if (spread_packing(ctx)) {
    std::vector<int4> lad;
    for (int c = 0; c < (int)ctx->cx.size(); ++c) {
      const HostComplex &h = ctx->cx[c];
      const int n_tmp = h.d.n_tmp, n_lig = h.d.n_lig;
      for (int p = 0; p < h.d.n_prt; ++p)
        for (int l = 0; l < n_lig; ++l)
          lad.push_back(make_int4(c, p * n_lig + l, h.rep_base + n_tmp * n_lig * p + l, n_lig));
    }
    ctx->n_ladders = (int)lad.size();
    HIPCHK(hipMalloc(&ctx->d_ladders, lad.size() * sizeof(int4)));
    HIPCHK(hipMemcpyAsync(ctx->d_ladders, lad.data(), lad.size() * sizeof(int4),
                          hipMemcpyHostToDevice, ctx->stream));
  }
  HIPCHK(hipMalloc(&ctx->d_packs, ps.packs.size() * sizeof(Pack)));
  HIPCHK(hipMalloc(&ctx->d_slots, ps.slots.size() * sizeof(int)));
  HIPCHK(hipMemcpyAsync(ctx->d_packs, ps.packs.data(), ps.packs.size() * sizeof(Pack),
                        hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(hipMemcpyAsync(ctx->d_slots, ps.slots.data(), ps.slots.size() * sizeof(int),
                        hipMemcpyHostToDevice, ctx->stream));
  ctx->n_packs = (int)ps.packs.size();
  ctx->mc_classes = ps.classes;
  HIPCHK(hipMalloc(&ctx->d_ctrl, (1 + ps.classes.size()) * sizeof(int)));
  HIPCHK(hipMemsetAsync(ctx->d_ctrl, 0, (1 + ps.classes.size()) * sizeof(int), ctx->stream));
  for (const LaunchClass &lc : ps.classes) {
    if (lc.lds > ctx->max_lds)
      return fail(ctx, GD_ERR_CAPACITY, "launch needs %zu bytes of LDS, device has %zu", lc.lds,
                  ctx->max_lds);
    if (lc.block > 512)
      return fail(ctx, GD_ERR_CAPACITY, "a temperature ladder needs %d lanes (> 512)", lc.block);
  }
  HIPCHK(hipStreamSynchronize(ctx->stream));
  ctx->state_alloc = true;
  if (ctx->opts.verbose) {
    fprintf(stderr, "[geauxdock] %d replicas, %d packs, split %s, %zu launch classes\n", n,
            ctx->n_packs, ns ? "fixed" : "auto", ctx->mc_classes.size());
    for (const LaunchClass &lc : ctx->mc_classes) {
      int nsp[4] = {0, 0, 0, 0};
      for (int i = lc.begin; i < lc.end; ++i) {
        const int v = ps.packs[i].ns;
        ++nsp[v >= 8 ? 3 : v >= 4 ? 2 : v >= 2 ? 1 : 0];
      }
      fprintf(stderr,
              "[geauxdock]   apl %d packs %d block %d LDS %zu B tile %d (split 1/2/4/8: %d/%d/%d/%d)\n",
              lc.apl, lc.end - lc.begin, lc.block, lc.lds, lc.tile, nsp[0], nsp[1], nsp[2],
              nsp[3]);
    }
  }
  return GD_OK;
}

KParams base_params(gd_ctx *ctx) {
  KParams P{};
  P.cxs = ctx->d_cx;
  P.packs = ctx->d_packs;
  P.slots = ctx->d_slots;
  P.mm = ctx->d_mm;
  P.e_old = ctx->d_eold;
  P.best_e = ctx->d_best_e;
  P.best_mm = ctx->d_best_mm;
  P.idx_tmp = ctx->d_idx_tmp;
  P.acs_mc = ctx->d_acs_mc;
  P.acs_x = ctx->d_acs_x;
  P.rep_flat = ctx->d_rep_flat;
  P.rec = ctx->d_ring[ctx->ring];
  P.rec_cnt = ctx->d_ring_cnt[ctx->ring];
  P.rec_cap = ctx->rec_cap;
  P.k0 = (uint32_t)ctx->opts.seed;
  P.k1 = (uint32_t)(ctx->opts.seed >> 32);
  return P;
}

// Kernel instance for atoms per lane `apl` and mode `da`: 0 plain,
// 1 delayed acceptance with the network per atom, 2 with its tables, 3 and 4
// the same with the split groups' shared surrogate sums compiled in.
const void *mc_kernel_fn(int apl, int da, bool tiled) {
  switch (apl) {
    case 1: return mc_kernel_handle<1>(da, tiled);
    case 2: return mc_kernel_handle<2>(da, tiled);
    case 3: return mc_kernel_handle<3>(da, tiled);
    default: return mc_kernel_handle<4>(da, tiled);
  }
}

// Instance mode of a delayed-acceptance or surrogate launch (see
// mc_kernel_fn): the shared-sum variants only when some complex's KDE or MCS
// sets are large enough to share, so other workloads run code without them.
int da_instance(const gd_ctx *ctx, bool tab) {
  bool any = false;
  for (const HostComplex &h : ctx->cx) any = any || h.dev.da_share != 0;
  return (tab ? 2 : 1) + (any ? 2 : 0);
}

// Launch every class of a pack set, the classes concurrently on side streams
// forked from and joined to `stream`. With `da`, the delayed-acceptance
// instances run; `da_scratch` adds the per-wave LDS scratch of the direct
// network evaluation (a class whose LDS would overflow then runs each pack as
// one lane group: the same packs and decisions' inputs, with the pair sums in
// one group's order).
// Per-class kernel arguments of a launch (persistent grids when `queue`).
struct ClassLaunch {
  const void *f;
  KParams P;
  int grid, block;
  size_t lds;
};

hipError_t plan_classes(gd_ctx *ctx, const std::vector<LaunchClass> &classes, const KParams &P0,
                        int da, bool da_scratch, int *queue, std::vector<ClassLaunch> &out) {
  out.clear();
  for (size_t i = 0; i < classes.size(); ++i) {
    const LaunchClass &lc = classes[i];
    ClassLaunch cl;
    cl.P = P0;
    KParams &P = cl.P;
    apply_layout(P, lc.lay);
    P.pack_base = lc.begin;
    cl.block = lc.block;
    cl.lds = lc.lds;
    if (da && da_scratch) {
      P.off_da = (int)((cl.lds + 15) & ~(size_t)15);
      cl.lds = P.off_da + (size_t)(cl.block / 64) * 6208;
      if (cl.lds > ctx->max_lds && lc.half_max < cl.block) {
        P.unsplit = 1;
        cl.block = lc.half_max;
        cl.lds = P.off_da + (size_t)(cl.block / 64) * 6208;
      }
      if (cl.lds > ctx->max_lds) return hipErrorInvalidValue;
    } else if (da) {
      P.off_da = (int)((cl.lds + 15) & ~(size_t)15);
    }
    const int n_pk = lc.end - lc.begin;
    cl.grid = n_pk;
    if (queue) {
      // resident workgroups: LDS-limited, and at most 8 waves per CU (the
      // kernels' register use allows two waves per SIMD)
      const int by_lds = (int)std::max<size_t>(1, ctx->max_lds / std::max<size_t>(cl.lds, 1));
      const int by_waves = std::max(1, 8 / std::max(1, cl.block / 64));
      cl.grid = std::min(n_pk, ctx->n_cu * std::min(by_lds, by_waves));
      P.queue = queue + i;
      P.n_queue = n_pk;
    }
    cl.f = mc_kernel_fn(lc.apl, da, lc.tile > 0);
    out.push_back(cl);
  }
  // the dynamic LDS limit of every instance used is the device's: classes
  // (and pose-evaluation launches) of different sizes share instances
  for (const ClassLaunch &cl : out) {
    const hipError_t e =
        hipFuncSetAttribute(cl.f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)ctx->max_lds);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

// Enqueue planned class launches, the classes concurrently on side streams
// forked from and joined to `stream` (capturable).
hipError_t enqueue_classes(gd_ctx *ctx, std::vector<ClassLaunch> &cls, int *queue,
                           hipStream_t stream) {
  const int n = (int)cls.size();
  hipError_t e = hipSuccess;
  if (queue && n > 0) {
    hipLaunchKernelGGL(gd_queue_reset_kernel, dim3((n + 63) / 64), dim3(64), 0, stream, queue, n);
    e = hipGetLastError();
  }
  const int n_side = std::min(n - 1, 3);
  if (e == hipSuccess && n_side > 0) {
    e = hipEventRecord(ctx->ev_fork, stream);
    for (int i = 0; i < n_side && e == hipSuccess; ++i) e = hipStreamWaitEvent(ctx->aux[i], ctx->ev_fork, 0);
  }
  for (int i = 0; i < n && e == hipSuccess; ++i) {
    void *args[] = {&cls[i].P};
    const int side = i % 4;  // 0: the main stream
    hipStream_t s = side == 0 ? stream : ctx->aux[side - 1];
    e = hipLaunchKernel(cls[i].f, dim3(cls[i].grid), dim3(cls[i].block), args, cls[i].lds, s);
  }
  for (int i = 0; i < n_side && e == hipSuccess; ++i) {
    e = hipEventRecord(ctx->ev_join[i], ctx->aux[i]);
    if (e == hipSuccess) e = hipStreamWaitEvent(stream, ctx->ev_join[i], 0);
  }
  return e;
}

hipError_t launch_classes(gd_ctx *ctx, const std::vector<LaunchClass> &classes, const KParams &P0,
                          int da, bool da_scratch, hipStream_t stream) {
  std::vector<ClassLaunch> cls;
  hipError_t e = plan_classes(ctx, classes, P0, da, da_scratch, nullptr, cls);
  if (e == hipSuccess) e = enqueue_classes(ctx, cls, nullptr, stream);
  return e;
}

// Drain record ring r: its accepted-state records, compacted on the device
// into one window, copied to the pinned staging buffer and handed to the
// sink per complex; then the ring's counts restart. All on `stream`, which
// is synchronized before returning.
gd_status drain_ring(gd_ctx *ctx, int r, gd_record_sink sink, void *user, hipStream_t stream) {
  if (!ctx->d_ring[r]) return GD_OK;
  const int n = ctx->n_rep;
  int *cnt_d = ctx->d_ring_cnt[r];
  if (sink) {
    std::vector<int> cnt(n);
    HIPCHK(hipMemcpyAsync(cnt.data(), cnt_d, n * sizeof(int), hipMemcpyDeviceToHost, stream));
    HIPCHK(hipStreamSynchronize(stream));
    std::vector<int64_t> off(n + 1, 0);
    for (int i = 0; i < n; ++i) off[i + 1] = off[i] + cnt[i];
    const int64_t tot = off[n];
    if (tot > 0) {
      if (tot > ctx->drain_cap) {
        if (ctx->d_drain) (void)hipFree(ctx->d_drain);
        if (ctx->h_drain) (void)hipHostFree(ctx->h_drain);
        ctx->d_drain = nullptr;
        ctx->h_drain = nullptr;
        ctx->drain_cap = 0;
        const int64_t cap = tot + tot / 8;
        HIPCHK(hipMalloc(&ctx->d_drain, cap * sizeof(gd_record)));
        HIPCHK(hipHostMalloc(&ctx->h_drain, cap * sizeof(gd_record), hipHostMallocDefault));
        ctx->drain_cap = cap;
      }
      if (!ctx->d_drain_off) HIPCHK(hipMalloc(&ctx->d_drain_off, (n + 1) * sizeof(int64_t)));
      HIPCHK(hipMemcpyAsync(ctx->d_drain_off, off.data(), (n + 1) * sizeof(int64_t),
                            hipMemcpyHostToDevice, stream));
      hipLaunchKernelGGL(gd_compact_kernel, dim3(n), dim3(256), 0, stream, n, ctx->d_ring[r],
                         cnt_d, ctx->d_drain_off, ctx->rec_cap, ctx->d_drain);
      HIPCHK(hipGetLastError());
      HIPCHK(hipMemcpyAsync(ctx->h_drain, ctx->d_drain, tot * sizeof(gd_record),
                            hipMemcpyDeviceToHost, stream));
      HIPCHK(hipStreamSynchronize(stream));
      for (int c = 0; c < (int)ctx->cx.size(); ++c) {
        const HostComplex &h = ctx->cx[c];
        const int64_t b = off[h.rep_base], e = off[h.rep_base + h.n_rep];
        if (e > b) sink(user, c, ctx->h_drain + b, e - b);
      }
    }
  }
  HIPCHK(hipMemsetAsync(cnt_d, 0, n * sizeof(int), stream));
  HIPCHK(hipStreamSynchronize(stream));
  return GD_OK;
}

gd_status drain_records(gd_ctx *ctx, gd_record_sink sink, void *user) {
  return drain_ring(ctx, ctx->ring, sink, user, ctx->stream);
}

// Second record ring, allocated the first time a run drains to a sink.
gd_status ensure_ring1(gd_ctx *ctx) {
  if (ctx->d_ring[1] || !ctx->d_rec) return GD_OK;
  const int n = ctx->n_rep;
  HIPCHK(hipMalloc(&ctx->d_ring[1], (size_t)n * ctx->rec_cap * sizeof(gd_record)));
  HIPCHK(hipMalloc(&ctx->d_ring_cnt[1], (size_t)n * sizeof(int)));
  HIPCHK(hipMemsetAsync(ctx->d_ring_cnt[1], 0, (size_t)n * sizeof(int), ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return GD_OK;
}

// Enqueue one MC launch (mode, window of n_steps from step0) on ctx->stream
// between the timing events ev0 and ev1; launch_mc_wait waits for it.
gd_status launch_mc_enqueue(gd_ctx *ctx, int mode, int step0, int n_steps, int nc, int da_k = 0,
                            float sur_scale = 1.0f) {
  KParams P = base_params(ctx);
  P.mode = mode;
  P.step0 = step0;
  P.n_steps = n_steps;
  P.nc = nc;
  bool scratch = false;
  if (da_k > 0 && mode == MODE_MC) {
    auto &S = ctx->sur;
    P.da_k = da_k;
    P.sur_scale = sur_scale;
    P.sW1T = S.W1T;
    P.sW2T = S.W2T;
    P.sb2 = S.p + SP_B2;
    P.sw3 = S.p + SP_W3;
    P.sb3 = S.p + SP_B3;
    P.sur_tab = S.tab_valid ? 1 : 0;
    scratch = !P.sur_tab;  // the direct network evaluation needs per-wave LDS scratch
  }
  const int da = P.da_k > 0 ? da_instance(ctx, P.sur_tab != 0) : 0;
  const bool persistent = (ctx->opts.launch & GD_LAUNCH_PERSISTENT) && mode == MODE_MC;
  const bool use_graph = !(ctx->opts.launch & GD_LAUNCH_NO_GRAPH) && mode == MODE_MC;
  int *queue = persistent ? ctx->d_ctrl + 1 : nullptr;
  if (use_graph) {
    // the first step comes from device memory so one graph serves every window
    P.step_dev = ctx->d_ctrl;
    // one graph slot per record ring (the ring's pointers are kernel arguments)
    auto &G = ctx->gslot[ctx->ring];
    const bool same = G.valid && G.mode == mode && G.win == n_steps && G.nc == nc &&
                      G.da == P.da_k && G.tab == P.sur_tab && G.scale == P.sur_scale &&
                      G.persistent == (int)persistent;
    if (!same) {
      if (G.exec) (void)hipGraphExecDestroy(G.exec);
      G.exec = nullptr;
      G.valid = false;
      std::vector<ClassLaunch> cls;
      hipError_t e = plan_classes(ctx, ctx->mc_classes, P, da, scratch, queue, cls);
      if (e == hipErrorInvalidValue)
        return fail(ctx, GD_ERR_CAPACITY, "delayed acceptance scratch does not fit in LDS");
      HIPCHK(e);
      hipGraph_t g = nullptr;
      HIPCHK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
      e = enqueue_classes(ctx, cls, queue, ctx->stream);
      const hipError_t e2 = hipStreamEndCapture(ctx->stream, &g);
      HIPCHK(e);
      HIPCHK(e2);
      e = hipGraphInstantiate(&G.exec, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      HIPCHK(e);
      G.valid = true;
      G.mode = mode;
      G.win = n_steps;
      G.nc = nc;
      G.da = P.da_k;
      G.tab = P.sur_tab;
      G.scale = P.sur_scale;
      G.persistent = (int)persistent;
    }
    ctx->h_step = step0;
    HIPCHK(hipMemcpyAsync(ctx->d_ctrl, &ctx->h_step, sizeof(int), hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipEventRecord(ctx->ev0, ctx->stream));
    HIPCHK(hipGraphLaunch(G.exec, ctx->stream));
  } else {
    std::vector<ClassLaunch> cls;
    hipError_t e = plan_classes(ctx, ctx->mc_classes, P, da, scratch, queue, cls);
    if (e == hipErrorInvalidValue)
      return fail(ctx, GD_ERR_CAPACITY, "delayed acceptance scratch does not fit in LDS");
    HIPCHK(e);
    HIPCHK(hipEventRecord(ctx->ev0, ctx->stream));
    HIPCHK(enqueue_classes(ctx, cls, queue, ctx->stream));
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventRecord(ctx->ev1, ctx->stream));
  return GD_OK;
}

gd_status launch_mc_wait(gd_ctx *ctx, float *ms) {
  HIPCHK(hipEventSynchronize(ctx->ev1));
  if (ms) HIPCHK(hipEventElapsedTime(ms, ctx->ev0, ctx->ev1));
  return GD_OK;
}

gd_status launch_mc(gd_ctx *ctx, int mode, int step0, int n_steps, int nc, float *ms,
                    int da_k = 0, float sur_scale = 1.0f) {
  TRY(launch_mc_enqueue(ctx, mode, step0, n_steps, nc, da_k, sur_scale));
  return launch_mc_wait(ctx, ms);
}

// One window of spread packing: MC launches that end on exchange steps, each
// followed by the ladder exchange kernel (timed with its launch). Advances
// ctx->step by each launch's steps as it completes.
gd_status spread_window(gd_ctx *ctx, int win, int nc, int da_k, float sur_scale, double *ms_sum,
                        int *launches) {
  int done = 0;
  while (done < win) {
    const int s = (int)ctx->step;
    int sub = win - done;
    if (nc > 0) sub = std::min(sub, nc - s % nc);
    TRY(launch_mc_enqueue(ctx, MODE_MC, s, sub, 0, da_k, sur_scale));
    if (nc > 0 && (s + sub) % nc == 0) {
      const KParams P = base_params(ctx);
      hipLaunchKernelGGL(gd_ladder_exchange_kernel, dim3((ctx->n_ladders + 63) / 64), dim3(64),
                         0, ctx->stream, ctx->d_cx, ctx->d_ladders, ctx->n_ladders, s + sub, nc,
                         P.k0, P.k1, ctx->d_eold, ctx->d_idx_tmp, ctx->d_acs_x);
      HIPCHK(hipGetLastError());
      HIPCHK(hipEventRecord(ctx->ev1, ctx->stream));
    }
    float ms = 0.f;
    TRY(launch_mc_wait(ctx, &ms));
    *ms_sum += ms;
    *launches += 1;
    ctx->step += sub;
    done += sub;
  }
  return GD_OK;
}

}  // namespace

extern "C" {

static gd_status run_impl(gd_ctx *ctx, int64_t n_steps, int steps_per_exchange,
                          gd_record_sink sink, void *user, gd_run_stats *stats, int da_k,
                          float sur_scale);

gd_status gd_run(gd_ctx *ctx, int64_t n_steps, int steps_per_exchange, gd_record_sink sink,
                 void *user, gd_run_stats *stats) {
  return run_impl(ctx, n_steps, steps_per_exchange, sink, user, stats, 0, 1.0f);
}

gd_status gd_run_da(gd_ctx *ctx, int64_t n_steps, int steps_per_exchange, int k_sub,
                    float sur_scale, gd_record_sink sink, void *user, gd_run_stats *stats) {
  if (!ctx) return GD_ERR_INVALID;
  if (k_sub < 1) return fail(ctx, GD_ERR_INVALID, "k_sub must be >= 1");
  if (ctx->sur.g.size() != ctx->cx.size()) return fail(ctx, GD_ERR_STATE, "call gd_sur_build_grids first");
  if (!ctx->sur.have_params) return fail(ctx, GD_ERR_STATE, "no surrogate parameters");
  if (n_steps % k_sub != 0 || (steps_per_exchange > 0 && steps_per_exchange % k_sub != 0) ||
      ctx->step % k_sub != 0)
    return fail(ctx, GD_ERR_INVALID, "step counts must be multiples of k_sub");
  return run_impl(ctx, n_steps, steps_per_exchange, sink, user, stats, k_sub, sur_scale);
}

static gd_status run_impl(gd_ctx *ctx, int64_t n_steps, int steps_per_exchange,
                          gd_record_sink sink, void *user, gd_run_stats *stats, int da_k,
                          float sur_scale) {
  if (!ctx) return GD_ERR_INVALID;
  if (n_steps < 0 || steps_per_exchange < 0) return fail(ctx, GD_ERR_INVALID, "negative step count");
  if (ctx->step + n_steps > 2000000000LL) return fail(ctx, GD_ERR_CAPACITY, "step counter overflow");
  HIPCHK(hipSetDevice(ctx->device));
  const auto t_start = std::chrono::steady_clock::now();
  TRY(alloc_state(ctx));
  gd_run_stats st{};
  std::vector<int> acs0(ctx->n_rep), acx0(ctx->n_rep);
  HIPCHK(hipMemcpyAsync(acs0.data(), ctx->d_acs_mc, ctx->n_rep * sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipMemcpyAsync(acx0.data(), ctx->d_acs_x, ctx->n_rep * sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  if (!ctx->initialized) {
    float ms = 0.f;
    TRY(launch_mc(ctx, MODE_INIT, (int)ctx->step, 0, 0, &ms));
    ctx->initialized = true;
  }
  int64_t left = n_steps;
  double t_ms = 0.0;
  // with records and a sink, windows alternate between the two rings: the
  // previous window's records are drained (on the drain stream, sinks called
  // here) while the next window runs
  const bool spread = spread_packing(ctx);
  const bool overlap = sink && ctx->d_rec && n_steps > 0 && !spread;
  if (overlap) TRY(ensure_ring1(ctx));
  int pending = -1;
  while (left > 0) {
    int win = (int)std::min<int64_t>(left, ctx->opts.steps_per_dump);
    if (da_k > 0) win = std::max(da_k, win / da_k * da_k);
    if (spread) {
      TRY(spread_window(ctx, win, steps_per_exchange, da_k, sur_scale, &t_ms, &st.kernel_launches));
      left -= win;
      TRY(drain_records(ctx, sink, user));
      continue;
    }
    float ms = 0.f;
    TRY(launch_mc_enqueue(ctx, MODE_MC, (int)ctx->step, win, steps_per_exchange, da_k, sur_scale));
    // a failed drain is reported after the window in flight has finished
    const gd_status ds =
        pending >= 0 ? drain_ring(ctx, pending, sink, user, ctx->drain_stream) : GD_OK;
    TRY(launch_mc_wait(ctx, &ms));
    t_ms += ms;
    st.kernel_launches += 1;
    ctx->step += win;
    left -= win;
    if (ds != GD_OK) return ds;
    if (overlap) {
      pending = ctx->ring;
      ctx->ring ^= 1;
    } else {
      TRY(drain_records(ctx, sink, user));
    }
  }
  if (pending >= 0) TRY(drain_ring(ctx, pending, sink, user, ctx->stream));
  if (n_steps == 0) TRY(drain_records(ctx, sink, user));
  std::vector<int> acs1(ctx->n_rep), acx1(ctx->n_rep);
  HIPCHK(hipMemcpyAsync(acs1.data(), ctx->d_acs_mc, ctx->n_rep * sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipMemcpyAsync(acx1.data(), ctx->d_acs_x, ctx->n_rep * sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  for (int i = 0; i < ctx->n_rep; ++i) {
    st.accepted += acs1[i] - acs0[i];
    st.exchanges_accepted += acx1[i] - acx0[i];
  }
  st.t_compute = t_ms * 1e-3;
  st.steps = n_steps;
  st.replica_steps = n_steps * (int64_t)ctx->n_rep;
  st.kernel_ms_avg = st.kernel_launches ? t_ms / st.kernel_launches : 0.0;
  st.t_wall = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
  if (stats) *stats = st;
  return GD_OK;
}

gd_status gd_get_state(gd_ctx *ctx, int cid, float *movematrix, gd_energy *energy_old,
                       gd_replica *replica, int32_t *acs_mc, int32_t *acs_x) {
  if (!ctx || cid < 0 || cid >= (int)ctx->cx.size()) return GD_ERR_INVALID;
  if (!ctx->state_alloc) return fail(ctx, GD_ERR_STATE, "no state before gd_run");
  HIPCHK(hipSetDevice(ctx->device));
  const HostComplex &h = ctx->cx[cid];
  const size_t b = h.rep_base, n = h.n_rep;
  if (movematrix)
    HIPCHK(hipMemcpyAsync(movematrix, ctx->d_mm + b * 6, n * 6 * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
  if (energy_old)
    HIPCHK(hipMemcpyAsync(energy_old, ctx->d_eold + b, n * sizeof(gd_energy), hipMemcpyDeviceToHost, ctx->stream));
  if (acs_mc) HIPCHK(hipMemcpyAsync(acs_mc, ctx->d_acs_mc + b, n * sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  if (acs_x) HIPCHK(hipMemcpyAsync(acs_x, ctx->d_acs_x + b, n * sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  std::vector<int> it(n);
  HIPCHK(hipMemcpyAsync(it.data(), ctx->d_idx_tmp + b, n * sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  if (replica) {
    for (size_t f = 0; f < n; ++f) {
      replica[f].idx_rep = (int)f;
      replica[f].idx_prt = (int)(f / (h.d.n_tmp * h.d.n_lig));
      replica[f].idx_tmp = it[f];
      replica[f].idx_lig = (int)(f % h.d.n_lig);
    }
  }
  return GD_OK;
}

gd_status gd_get_best(gd_ctx *ctx, int cid, float *best_e, float *best_mm) {
  if (!ctx || cid < 0 || cid >= (int)ctx->cx.size()) return GD_ERR_INVALID;
  if (!ctx->initialized) return fail(ctx, GD_ERR_STATE, "no state before gd_run");
  HIPCHK(hipSetDevice(ctx->device));
  const HostComplex &h = ctx->cx[cid];
  const size_t b = h.rep_base, n = h.n_rep;
  if (best_e)
    HIPCHK(hipMemcpyAsync(best_e, ctx->d_best_e + b, n * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
  if (best_mm)
    HIPCHK(hipMemcpyAsync(best_mm, ctx->d_best_mm + b * 6, n * 6 * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return GD_OK;
}

gd_status gd_reset_best(gd_ctx *ctx) {
  if (!ctx) return GD_ERR_INVALID;
  if (!ctx->initialized) return fail(ctx, GD_ERR_STATE, "no state before gd_run");
  HIPCHK(hipSetDevice(ctx->device));
  const int n = ctx->n_rep;
  hipLaunchKernelGGL(gd_reset_best_kernel, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, n,
                     ctx->d_eold, ctx->d_mm, ctx->d_best_e, ctx->d_best_mm);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return GD_OK;
}

static const uint32_t kCkptMagic = 0x47444b32u;  // "GDK2"

gd_status gd_checkpoint(gd_ctx *ctx, void *buf, size_t *size) {
  if (!ctx || !size) return GD_ERR_INVALID;
  if (!ctx->state_alloc) return fail(ctx, GD_ERR_STATE, "no state before gd_run");
  const size_t n = ctx->n_rep;
  const size_t need = 32 + n * (13 * sizeof(float) + sizeof(gd_energy) + 3 * sizeof(int));
  if (!buf || *size < need) {
    *size = need;
    return buf ? fail(ctx, GD_ERR_INVALID, "checkpoint buffer too small") : GD_OK;
  }
  HIPCHK(hipSetDevice(ctx->device));
  char *o = (char *)buf;
  uint32_t hdr[8] = {kCkptMagic, (uint32_t)n, (uint32_t)ctx->initialized, 0,
                     (uint32_t)(ctx->step & 0xffffffffu), (uint32_t)(ctx->step >> 32),
                     (uint32_t)ctx->cx.size(), 0};
  memcpy(o, hdr, 32);
  o += 32;
  HIPCHK(hipMemcpyAsync(o, ctx->d_mm, n * 6 * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
  o += n * 6 * sizeof(float);
  HIPCHK(hipMemcpyAsync(o, ctx->d_eold, n * sizeof(gd_energy), hipMemcpyDeviceToHost, ctx->stream));
  o += n * sizeof(gd_energy);
  HIPCHK(hipMemcpyAsync(o, ctx->d_idx_tmp, n * sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  o += n * sizeof(int);
  HIPCHK(hipMemcpyAsync(o, ctx->d_acs_mc, n * sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  o += n * sizeof(int);
  HIPCHK(hipMemcpyAsync(o, ctx->d_acs_x, n * sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  o += n * sizeof(int);
  HIPCHK(hipMemcpyAsync(o, ctx->d_best_e, n * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
  o += n * sizeof(float);
  HIPCHK(hipMemcpyAsync(o, ctx->d_best_mm, n * 6 * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  *size = need;
  return GD_OK;
}

gd_status gd_restore(gd_ctx *ctx, const void *buf, size_t size) {
  if (!ctx || !buf) return GD_ERR_INVALID;
  HIPCHK(hipSetDevice(ctx->device));
  TRY(alloc_state(ctx));
  const size_t n = ctx->n_rep;
  const size_t need = 32 + n * (13 * sizeof(float) + sizeof(gd_energy) + 3 * sizeof(int));
  uint32_t hdr[8];
  if (size < need) return fail(ctx, GD_ERR_INVALID, "checkpoint too small");
  memcpy(hdr, buf, 32);
  if (hdr[0] != kCkptMagic || hdr[1] != n || hdr[6] != ctx->cx.size())
    return fail(ctx, GD_ERR_INVALID, "checkpoint does not match this context");
  const char *o = (const char *)buf + 32;
  HIPCHK(hipMemcpyAsync(ctx->d_mm, o, n * 6 * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
  o += n * 6 * sizeof(float);
  HIPCHK(hipMemcpyAsync(ctx->d_eold, o, n * sizeof(gd_energy), hipMemcpyHostToDevice, ctx->stream));
  o += n * sizeof(gd_energy);
  HIPCHK(hipMemcpyAsync(ctx->d_idx_tmp, o, n * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
  o += n * sizeof(int);
  HIPCHK(hipMemcpyAsync(ctx->d_acs_mc, o, n * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
  o += n * sizeof(int);
  HIPCHK(hipMemcpyAsync(ctx->d_acs_x, o, n * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
  o += n * sizeof(int);
  HIPCHK(hipMemcpyAsync(ctx->d_best_e, o, n * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
  o += n * sizeof(float);
  HIPCHK(hipMemcpyAsync(ctx->d_best_mm, o, n * 6 * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
  for (int r = 0; r < 2; ++r)
    if (ctx->d_ring_cnt[r]) HIPCHK(hipMemsetAsync(ctx->d_ring_cnt[r], 0, n * sizeof(int), ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  ctx->initialized = hdr[2] != 0;
  ctx->step = (int64_t)hdr[4] | ((int64_t)hdr[5] << 32);
  return GD_OK;
}

static gd_status eval_poses(gd_ctx *ctx, int cid, int n, const int *idx_lig, const int *idx_prt,
                            const float *movematrix, gd_energy *out, float *out_sur,
                            float sur_scale, bool no_alloc = false);

gd_status gd_calc_energy_noalloc(gd_ctx *ctx, int cid, int n, const int *idx_lig,
                                 const int *idx_prt, const float *movematrix, gd_energy *out) {
  if (!ctx || cid < 0 || cid >= (int)ctx->cx.size() || n < 0) return GD_ERR_INVALID;
  if (n == 0) return GD_OK;
  if (!idx_lig || !idx_prt || !movematrix || !out) return GD_ERR_INVALID;
  if (n > GD_NOALLOC_MAX_POSES)
    return fail(ctx, GD_ERR_CAPACITY, "gd_calc_energy_noalloc: %d poses > %d per call", n,
                GD_NOALLOC_MAX_POSES);
  if (!ctx->scratch) return fail(ctx, GD_ERR_STATE, "no device scratch address");
  if (ctx->cx_dirty || !ctx->cx[cid].ref_built)
    return fail(ctx, GD_ERR_STATE,
                "gd_calc_energy_noalloc: the complex table and contact map are not on the device "
                "yet; call gd_calc_energy or gd_run once after the last gd_add_complex");
  return eval_poses(ctx, cid, n, idx_lig, idx_prt, movematrix, out, nullptr, 0.0f, true);
}

gd_status gd_calc_energy(gd_ctx *ctx, int cid, int n, const int *idx_lig, const int *idx_prt,
                         const float *movematrix, gd_energy *out) {
  if (!ctx || cid < 0 || cid >= (int)ctx->cx.size() || n < 0) return GD_ERR_INVALID;
  if (n == 0) return GD_OK;
  if (!idx_lig || !idx_prt || !movematrix || !out) return GD_ERR_INVALID;
  return eval_poses(ctx, cid, n, idx_lig, idx_prt, movematrix, out, nullptr, 0.0f);
}

gd_status gd_sur_surrogate(gd_ctx *ctx, int cid, int n, const int *idx_lig, const int *idx_prt,
                           const float *movematrix, float sur_scale, float *out) {
  if (!ctx || cid < 0 || cid >= (int)ctx->cx.size() || n < 0) return GD_ERR_INVALID;
  if (n == 0) return GD_OK;
  if (!idx_lig || !idx_prt || !movematrix || !out) return GD_ERR_INVALID;
  if (ctx->sur.g.size() != ctx->cx.size()) return fail(ctx, GD_ERR_STATE, "call gd_sur_build_grids first");
  if (!ctx->sur.have_params) return fail(ctx, GD_ERR_STATE, "no surrogate parameters");
  return eval_poses(ctx, cid, n, idx_lig, idx_prt, movematrix, nullptr, out, sur_scale);
}

// Energies (or surrogate totals) of n poses of one complex. With no_alloc the
// pose tables and outputs live in the device scratch array (the caller has
// checked n and that the complex table and contact map are on the device);
// otherwise they are allocated for the call.
static gd_status eval_poses(gd_ctx *ctx, int cid, int n, const int *idx_lig, const int *idx_prt,
                            const float *movematrix, gd_energy *out, float *out_sur,
                            float sur_scale, bool no_alloc) {
  HostComplex &h = ctx->cx[cid];
  for (int i = 0; i < n; ++i)
    if (idx_lig[i] < 0 || idx_lig[i] >= h.d.n_lig || idx_prt[i] < 0 || idx_prt[i] >= h.d.n_prt)
      return fail(ctx, GD_ERR_INVALID, "pose %d: conformer index out of range", i);
  HIPCHK(hipSetDevice(ctx->device));
  if (!no_alloc) TRY(ensure_ref(ctx));
  // each pose is evaluated with the receptor split of its ligand conformer's
  // MC packs, so a pose's energy has the bits an MC step gives it
  std::vector<int> lig_ns = h.lig_ns;
  if (lig_ns.size() != (size_t)h.d.n_lig) lig_ns = lig_splits(ctx, h);
  PackSet ps;
  for (int p = 0; p < h.d.n_prt; ++p)
    for (int ns = 1; ns <= 8; ns *= 2) {
      std::vector<int> mine;
      for (int i = 0; i < n; ++i)
        if (idx_prt[i] == p && lig_ns[idx_lig[i]] == ns) mine.push_back(i);
      const int per = std::max(1, (512 / ns) / h.lpr);
      for (size_t b = 0; b < mine.size(); b += per) {
        Pack pk{};
        pk.cx = cid;
        pk.prt = p;
        pk.n_slots = (int)std::min<size_t>(per, mine.size() - b);
        pk.n_groups = 0;
        pk.ns = ns;
        pk.slot_off = (int)ps.slots.size();
        for (int j = 0; j < pk.n_slots; ++j) ps.slots.push_back(mine[b + j]);
        ps.packs.push_back(pk);
      }
    }
  pack_finish(ctx, ps);
  for (const LaunchClass &lc : ps.classes)
    if (lc.lds > ctx->max_lds) return fail(ctx, GD_ERR_CAPACITY, "LDS %zu > %zu", lc.lds, ctx->max_lds);
  KParams P{};
  const bool sur = out_sur != nullptr;
  if (sur) {
    auto &S = ctx->sur;
    P.sur_scale = sur_scale;
    P.sur_tab = S.tab_valid ? 1 : 0;
    P.sW1T = S.W1T;
    P.sW2T = S.W2T;
    P.sb2 = S.p + SP_B2;
    P.sw3 = S.p + SP_W3;
    P.sb3 = S.p + SP_B3;
  }
  Pack *d_packs = nullptr;
  int *d_slots = nullptr, *d_lig = nullptr;
  float *d_mm = nullptr;
  gd_energy *d_out = nullptr;
  float *d_sur = nullptr;
  if (no_alloc) {
    // regions of the scratch array, each 256-byte aligned
    auto al = [](size_t v) { return (v + 255) & ~(size_t)255; };
    const size_t o_slots = al(ps.packs.size() * sizeof(Pack));
    const size_t o_lig = o_slots + al((size_t)n * sizeof(int));
    const size_t o_mm = o_lig + al((size_t)n * sizeof(int));
    const size_t o_out = o_mm + al((size_t)n * 6 * sizeof(float));
    const size_t end = o_out + al((size_t)n * sizeof(gd_energy));
    if (end > GD_NOALLOC_SCRATCH_BYTES)
      return fail(ctx, GD_ERR_CAPACITY, "gd_calc_energy_noalloc: %zu scratch bytes needed, %d held",
                  end, GD_NOALLOC_SCRATCH_BYTES);
    d_packs = reinterpret_cast<Pack *>(ctx->scratch);
    d_slots = reinterpret_cast<int *>(ctx->scratch + o_slots);
    d_lig = reinterpret_cast<int *>(ctx->scratch + o_lig);
    d_mm = reinterpret_cast<float *>(ctx->scratch + o_mm);
    d_out = reinterpret_cast<gd_energy *>(ctx->scratch + o_out);
  } else {
    HIPCHK(hipMalloc(&d_packs, ps.packs.size() * sizeof(Pack)));
    HIPCHK(hipMalloc(&d_slots, ps.slots.size() * sizeof(int)));
    HIPCHK(hipMalloc(&d_lig, n * sizeof(int)));
    HIPCHK(hipMalloc(&d_mm, (size_t)n * 6 * sizeof(float)));
    HIPCHK(hipMalloc(&d_out, (size_t)n * sizeof(gd_energy)));
    HIPCHK(hipMalloc(&d_sur, (size_t)n * sizeof(float)));
  }
  HIPCHK(hipMemcpyAsync(d_packs, ps.packs.data(), ps.packs.size() * sizeof(Pack), hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(hipMemcpyAsync(d_slots, ps.slots.data(), ps.slots.size() * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(hipMemcpyAsync(d_lig, idx_lig, n * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(hipMemcpyAsync(d_mm, movematrix, (size_t)n * 6 * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
  P.cxs = ctx->d_cx;
  P.packs = d_packs;
  P.slots = d_slots;
  P.ev_mm = d_mm;
  P.ev_lig = d_lig;
  P.ev_out = d_out;
  P.ev_sur = d_sur;
  P.mode = sur ? MODE_SUREVAL : MODE_EVAL;
  hipError_t le = launch_classes(ctx, ps.classes, P, sur ? da_instance(ctx, P.sur_tab != 0) : 0,
                                 sur && !P.sur_tab, ctx->stream);
  if (le == hipSuccess && out)
    le = hipMemcpyAsync(out, d_out, (size_t)n * sizeof(gd_energy), hipMemcpyDeviceToHost, ctx->stream);
  if (le == hipSuccess && sur)
    le = hipMemcpyAsync(out_sur, d_sur, (size_t)n * sizeof(float), hipMemcpyDeviceToHost, ctx->stream);
  if (le == hipSuccess) le = hipStreamSynchronize(ctx->stream);
  if (!no_alloc) {
    (void)hipFree(d_sur);
    (void)hipFree(d_packs);
    (void)hipFree(d_slots);
    (void)hipFree(d_lig);
    (void)hipFree(d_mm);
    (void)hipFree(d_out);
  }
  if (le != hipSuccess) return fail(ctx, GD_ERR_HIP, "energy evaluation: %s", hipGetErrorString(le));
  return GD_OK;
}

}  // extern "C"

namespace {
template <class T>
gd_status tmp_dev(gd_ctx *ctx, std::vector<void *> &keep, T **p, const T *src, size_t n) {
  HIPCHK(hipMalloc((void **)p, std::max<size_t>(n, 1) * sizeof(T)));
  keep.push_back(*p);
  if (src && n) HIPCHK(hipMemcpyAsync(*p, src, n * sizeof(T), hipMemcpyHostToDevice, ctx->stream));
  return GD_OK;
}

void free_all(std::vector<void *> &keep) {
  for (void *p : keep) (void)hipFree(p);
  keep.clear();
}
}  // namespace

extern "C" {

gd_status gd_accept(gd_ctx *ctx, int n, const float *e_new, const float *e_old,
                    const float *minus_beta, const float *u, int32_t *accept) {
  if (!ctx || n < 0) return GD_ERR_INVALID;
  if (n == 0) return GD_OK;
  if (!e_new || !e_old || !minus_beta || !u || !accept) return GD_ERR_INVALID;
  HIPCHK(hipSetDevice(ctx->device));
  std::vector<void *> keep;
  float *a, *b, *c, *d;
  int *r;
  gd_status s = GD_OK;
  if ((s = tmp_dev(ctx, keep, &a, e_new, n)) || (s = tmp_dev(ctx, keep, &b, e_old, n)) ||
      (s = tmp_dev(ctx, keep, &c, minus_beta, n)) || (s = tmp_dev(ctx, keep, &d, u, n)) ||
      (s = tmp_dev<int>(ctx, keep, &r, nullptr, n))) {
    free_all(keep);
    return s;
  }
  hipLaunchKernelGGL(gd_accept_kernel, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, n, a, b, c, d, r);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(accept, r, n * sizeof(int), hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  free_all(keep);
  if (e != hipSuccess) return fail(ctx, GD_ERR_HIP, "gd_accept: %s", hipGetErrorString(e));
  return GD_OK;
}

gd_status gd_exchange(gd_ctx *ctx, int n_groups, int n_tmp, int mode_t, const float *etot,
                      const float *minus_beta, const float *u, int32_t *idx_tmp, int32_t *acs) {
  if (!ctx || n_groups < 0 || n_tmp <= 0) return GD_ERR_INVALID;
  if (n_tmp > GD_MAX_NTMP) return fail(ctx, GD_ERR_CAPACITY, "n_tmp %d > %d", n_tmp, GD_MAX_NTMP);
  if (mode_t != 0 && mode_t != 1) return fail(ctx, GD_ERR_INVALID, "mode_t must be 0 or 1");
  if (n_groups == 0) return GD_OK;
  if (!etot || !minus_beta || !u || !idx_tmp || !acs) return GD_ERR_INVALID;
  const size_t ns = (size_t)n_groups * n_tmp;
  for (size_t i = 0; i < ns; ++i)
    if (idx_tmp[i] < 0 || idx_tmp[i] >= n_tmp)
      return fail(ctx, GD_ERR_INVALID, "idx_tmp[%zu] = %d outside [0,%d)", i, idx_tmp[i], n_tmp);
  HIPCHK(hipSetDevice(ctx->device));
  std::vector<void *> keep;
  float *de, *dm, *du;
  int *di, *da;
  gd_status s = GD_OK;
  const size_t nu = (size_t)n_groups * (n_tmp / 2);
  if ((s = tmp_dev(ctx, keep, &de, etot, ns)) || (s = tmp_dev(ctx, keep, &dm, minus_beta, (size_t)n_tmp)) ||
      (s = tmp_dev(ctx, keep, &du, u, nu)) || (s = tmp_dev<int>(ctx, keep, &di, idx_tmp, ns)) ||
      (s = tmp_dev<int>(ctx, keep, &da, acs, ns))) {
    free_all(keep);
    return s;
  }
  hipLaunchKernelGGL(gd_exchange_kernel, dim3((n_groups + 63) / 64), dim3(64), 0, ctx->stream,
                     n_groups, n_tmp, mode_t, de, dm, du, di, da);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(idx_tmp, di, ns * sizeof(int), hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(acs, da, ns * sizeof(int), hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  free_all(keep);
  if (e != hipSuccess) return fail(ctx, GD_ERR_HIP, "gd_exchange: %s", hipGetErrorString(e));
  return GD_OK;
}

gd_status gd_philox(gd_ctx *ctx, int n, const uint32_t *ctr, const uint32_t *key, uint32_t *out) {
  if (!ctx || n < 0) return GD_ERR_INVALID;
  if (n == 0) return GD_OK;
  if (!ctr || !key || !out) return GD_ERR_INVALID;
  HIPCHK(hipSetDevice(ctx->device));
  std::vector<void *> keep;
  uint32_t *dc, *dk, *dout;
  gd_status s = GD_OK;
  if ((s = tmp_dev(ctx, keep, &dc, ctr, (size_t)4 * n)) || (s = tmp_dev(ctx, keep, &dk, key, (size_t)2 * n)) ||
      (s = tmp_dev<uint32_t>(ctx, keep, &dout, nullptr, (size_t)4 * n))) {
    free_all(keep);
    return s;
  }
  hipLaunchKernelGGL(gd_philox_kernel, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, n, dc, dk, dout);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(out, dout, (size_t)4 * n * sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  free_all(keep);
  if (e != hipSuccess) return fail(ctx, GD_ERR_HIP, "gd_philox: %s", hipGetErrorString(e));
  return GD_OK;
}

}  // extern "C"

// ===========================================================================
// Learned surrogate: feature grids, dataset, BF16 MFMA training, prediction.
// ===========================================================================
namespace {

gd_status sur_alloc_params(gd_ctx *ctx) {
  auto &S = ctx->sur;
  if (S.p) return GD_OK;
  HIPCHK(hipMalloc(&S.p, SUR_NPARAM * sizeof(float)));
  HIPCHK(hipMalloc(&S.m, SUR_NPARAM * sizeof(float)));
  HIPCHK(hipMalloc(&S.v, SUR_NPARAM * sizeof(float)));
  HIPCHK(hipMalloc(&S.grad, SUR_NPARAM * sizeof(float)));
  HIPCHK(hipMalloc(&S.W1T, SUR_F * SUR_H * 2));
  HIPCHK(hipMalloc(&S.W2T, SUR_H * SUR_H * 2));
  HIPCHK(hipMalloc(&S.W2R, SUR_H * SUR_H * 2));
  return GD_OK;
}

SurW sur_weights(gd_ctx *ctx) {
  auto &S = ctx->sur;
  SurW W;
  W.W1T = S.W1T;
  W.W2T = S.W2T;
  W.W2R = S.W2R;
  W.b2 = S.p + SP_B2;
  W.w3 = S.p + SP_W3;
  W.b3 = S.p + SP_B3;
  return W;
}

// refresh the bf16 layouts from the FP32 master (Adam with lr = 0)
gd_status sur_refresh(gd_ctx *ctx) {
  auto &S = ctx->sur;
  hipLaunchKernelGGL(gd_mlp_adam_kernel, dim3((SUR_NPARAM + 255) / 256), dim3(256), 0, ctx->stream,
                     S.p, S.m, S.v, S.grad, 0.0f, 0.f, 0.f, 0.f, 1.f, 1.f, S.W1T, S.W2T, S.W2R);
  HIPCHK(hipGetLastError());
  return GD_OK;
}

gd_status sur_upload_index(gd_ctx *ctx) {
  auto &S = ctx->sur;
  if (!S.dirty) return GD_OK;
  const int64_t np = (int64_t)S.target.size();
  if (S.d_off) (void)hipFree(S.d_off);
  if (S.d_target) (void)hipFree(S.d_target);
  if (S.pose_of_row) (void)hipFree(S.pose_of_row);
  S.d_off = nullptr;
  S.d_target = nullptr;
  S.pose_of_row = nullptr;
  HIPCHK(hipMalloc(&S.d_off, (np + 1) * sizeof(int64_t)));
  HIPCHK(hipMalloc(&S.d_target, std::max<int64_t>(np, 1) * sizeof(float)));
  HIPCHK(hipMalloc(&S.pose_of_row, (S.cap + 64) * sizeof(int)));
  // per-row outputs are indexed by absolute dataset row
  if (S.y) (void)hipFree(S.y);
  S.y = nullptr;
  HIPCHK(hipMalloc(&S.y, (S.cap + 64) * sizeof(float)));
  std::vector<int> por(S.cap + 64, -1);
  for (int64_t p = 0; p < np; ++p)
    for (int64_t r = S.off[p]; r < S.off[p + 1]; ++r) por[r] = (int)p;
  HIPCHK(hipMemcpyAsync(S.d_off, S.off.data(), (np + 1) * sizeof(int64_t), hipMemcpyHostToDevice, ctx->stream));
  if (np) HIPCHK(hipMemcpyAsync(S.d_target, S.target.data(), np * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(hipMemcpyAsync(S.pose_of_row, por.data(), por.size() * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  S.dirty = false;
  return GD_OK;
}

// Workgroups of the batch kernels: the fused backward runs one per CU (its
// LDS tiles take 124 KB), the forward three (its registers allow three waves
// per SIMD); each walks 64-row steps.
int sur_train_groups(const gd_ctx *ctx, int64_t rows) {
  return (int)std::max<int64_t>(1, std::min<int64_t>(rows / 64, ctx->n_cu));
}
int sur_fwd_groups(const gd_ctx *ctx, int64_t rows) {
  return (int)std::max<int64_t>(1, std::min<int64_t>(rows / 64, 3 * (int64_t)ctx->n_cu));
}

// work buffers for batches of up to `rows` rows (rounded up to 64): the
// gradient partials of the fused backward's workgroups and per-pose outputs
gd_status sur_work(gd_ctx *ctx, int64_t rows, int64_t n_pose_total) {
  auto &S = ctx->sur;
  rows = (rows + 63) / 64 * 64 + 64;
  if (rows > S.work_rows) {
    void *qs[] = {S.p1, S.p2, S.pb2, S.pw3, S.dyr};
    for (void *p : qs)
      if (p) (void)hipFree(p);
    const int64_t ng = std::max(1, ctx->n_cu);
    HIPCHK(hipMalloc(&S.dyr, (size_t)rows * sizeof(float)));
    HIPCHK(hipMalloc(&S.p1, (size_t)ng * SUR_F * SUR_H * sizeof(float)));
    HIPCHK(hipMalloc(&S.p2, (size_t)ng * SUR_H * SUR_H * sizeof(float)));
    HIPCHK(hipMalloc(&S.pb2, (size_t)ng * SUR_H * sizeof(float)));
    HIPCHK(hipMalloc(&S.pw3, (size_t)ng * (SUR_H + 1) * sizeof(float)));
    HIPCHK(hipFuncSetAttribute((const void *)gd_mlp_train_kernel,
                               hipFuncAttributeMaxDynamicSharedMemorySize, SUR_TRAIN_LDS));
    S.work_rows = rows;
  }
  if (S.dS) (void)hipFree(S.dS);
  if (S.sq) (void)hipFree(S.sq);
  if (S.S) (void)hipFree(S.S);
  const int64_t np = std::max<int64_t>(n_pose_total, 1);
  HIPCHK(hipMalloc(&S.dS, np * sizeof(float)));
  HIPCHK(hipMalloc(&S.sq, np * sizeof(float)));
  HIPCHK(hipMalloc(&S.S, np * sizeof(float)));
  return GD_OK;
}

// forward (+ backward and gradient when grad) over poses [p0, p1) of the dataset
// With `update`, the gradient partials are left for gd_mlp_update_kernel
// (the training step sums them and applies Adam in one launch).
gd_status sur_batch(gd_ctx *ctx, int64_t p0, int64_t p1, bool grad, double *loss,
                    bool update = false) {
  auto &S = ctx->sur;
  const int64_t r0 = S.off[p0], r1 = S.off[p1];
  const int64_t nr = (r1 - r0 + 63) / 64 * 64;  // padded rows read from the dataset
  const int np = (int)(p1 - p0);
  const SurW W = sur_weights(ctx);
  const unsigned short *X = S.X + r0 * SUR_F;
  hipLaunchKernelGGL(gd_mlp_fwd_kernel, dim3(sur_fwd_groups(ctx, nr)), dim3(256), 0, ctx->stream, X,
                     nr, W, S.y + r0);
  hipLaunchKernelGGL(gd_mlp_pose_kernel, dim3((np + 255) / 256), dim3(256), 0, ctx->stream, S.y,
                     S.d_off, (int)p0, (int)p1, S.d_target, S.dS, S.sq, S.S);
  if (grad) {
    const int ng = sur_train_groups(ctx, nr);
    hipLaunchKernelGGL(gd_mlp_dyrow_kernel, dim3((unsigned)((nr + 255) / 256)), dim3(256), 0,
                       ctx->stream, S.pose_of_row + r0, nr, (int)p0, (int)p1, S.dS, S.dyr);
    hipLaunchKernelGGL(gd_mlp_train_kernel, dim3(ng), dim3(256), SUR_TRAIN_LDS, ctx->stream, X,
                       nr, S.dyr, W, S.p1, S.p2, S.pb2, S.pw3);
    S.last_groups = ng;
    if (!update)
      hipLaunchKernelGGL(gd_mlp_gradsum_kernel, dim3((SUR_NPARAM + 255) / 256), dim3(256), 0,
                         ctx->stream, S.p1, ng, S.p2, ng, S.pb2, S.pw3, (int64_t)ng, S.grad);
  }
  HIPCHK(hipGetLastError());
  if (loss) {
    std::vector<float> sq(np);
    HIPCHK(hipMemcpyAsync(sq.data(), S.sq + p0, np * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    double s = 0.0;
    for (float v : sq) s += v;
    *loss = s / np;
  }
  return GD_OK;
}

}  // namespace

extern "C" {

int gd_sur_num_params(void) { return SUR_NPARAM; }

gd_status gd_sur_build_grids(gd_ctx *ctx, float spacing, float half_width) {
  if (!ctx || !(spacing > 0.0f) || !(half_width > spacing)) return GD_ERR_INVALID;
  HIPCHK(hipSetDevice(ctx->device));
  TRY(sync_complexes(ctx));
  auto &S = ctx->sur;
  const int n = (int)std::ceil(2.0f * half_width / spacing) + 1;
  const size_t node_bytes = 32 * sizeof(float);
  for (float *g : S.grid)
    if (g) (void)hipFree(g);
  for (float *g : S.tab)
    if (g) (void)hipFree(g);
  S.tab.assign(ctx->cx.size(), nullptr);
  S.tab_valid = false;
  S.grid.assign(ctx->cx.size(), nullptr);
  S.g.assign(ctx->cx.size(), SurGrid{});
  for (size_t c = 0; c < ctx->cx.size(); ++c) {
    const HostComplex &h = ctx->cx[c];
    const size_t nodes = (size_t)n * n * n;
    HIPCHK(hipMalloc(&S.grid[c], (size_t)h.d.n_prt * nodes * node_bytes));
    SurGrid G;
    G.grid = S.grid[c];
    G.ox = h.d.pocket_center[0] - half_width;
    G.oy = h.d.pocket_center[1] - half_width;
    G.oz = h.d.pocket_center[2] - half_width;
    G.h = spacing;
    G.n = n;
    S.g[c] = G;
    ctx->cx[c].dev.sgrid = G.grid;
    ctx->cx[c].dev.sox = G.ox;
    ctx->cx[c].dev.soy = G.oy;
    ctx->cx[c].dev.soz = G.oz;
    ctx->cx[c].dev.sh = G.h;
    ctx->cx[c].dev.sn = G.n;
    ctx->cx[c].dev.stab = nullptr;
    ctx->cx_dirty = true;
    for (int p = 0; p < h.d.n_prt; ++p) {
      hipLaunchKernelGGL(gd_sur_grid_kernel, dim3((unsigned)((nodes + 255) / 256)), dim3(256), 0,
                         ctx->stream, h.dev.prt_xyze + (size_t)p * h.d.pnp, h.dev.prt_cls, h.d.pnp,
                         G.ox, G.oy, G.oz, G.h, n, S.grid[c] + (size_t)p * nodes * 32);
      HIPCHK(hipGetLastError());
    }
  }
  HIPCHK(hipStreamSynchronize(ctx->stream));
  TRY(sync_complexes(ctx));
  return GD_OK;
}

gd_status gd_sur_features(gd_ctx *ctx, int cid, int n, const int *idx_lig, const int *idx_prt,
                          const float *movematrix, float *out) {
  if (!ctx || cid < 0 || cid >= (int)ctx->cx.size() || n < 0) return GD_ERR_INVALID;
  if ((int)ctx->sur.g.size() <= cid) return fail(ctx, GD_ERR_STATE, "call gd_sur_build_grids first");
  if (n == 0) return GD_OK;
  const HostComplex &h = ctx->cx[cid];
  for (int i = 0; i < n; ++i)
    if (idx_lig[i] < 0 || idx_lig[i] >= h.d.n_lig || idx_prt[i] < 0 || idx_prt[i] >= h.d.n_prt)
      return fail(ctx, GD_ERR_INVALID, "pose %d: conformer index out of range", i);
  HIPCHK(hipSetDevice(ctx->device));
  std::vector<void *> keep;
  int *dl, *dp;
  float *dm, *dout;
  gd_status s = GD_OK;
  const size_t rows = (size_t)n * h.d.lna;
  if ((s = tmp_dev(ctx, keep, &dl, idx_lig, (size_t)n)) || (s = tmp_dev(ctx, keep, &dp, idx_prt, (size_t)n)) ||
      (s = tmp_dev(ctx, keep, &dm, movematrix, (size_t)n * 6)) ||
      (s = tmp_dev<float>(ctx, keep, &dout, nullptr, rows * SUR_NF))) {
    free_all(keep);
    return s;
  }
  hipLaunchKernelGGL(gd_sur_feat_kernel, dim3((unsigned)((rows + 255) / 256)), dim3(256), 0,
                     ctx->stream, ctx->d_cx, cid, ctx->sur.g[cid], n, dl, dp, dm, (int64_t)0,
                     (unsigned short *)nullptr, (unsigned short *)nullptr, (int64_t)0, dout);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(out, dout, rows * SUR_NF * sizeof(float), hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  free_all(keep);
  if (e != hipSuccess) return fail(ctx, GD_ERR_HIP, "gd_sur_features: %s", hipGetErrorString(e));
  return GD_OK;
}

gd_status gd_sur_set_params(gd_ctx *ctx, const float *params) {
  if (!ctx || !params) return GD_ERR_INVALID;
  HIPCHK(hipSetDevice(ctx->device));
  TRY(sur_alloc_params(ctx));
  auto &S = ctx->sur;
  HIPCHK(hipMemcpyAsync(S.p, params, SUR_NPARAM * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(hipMemsetAsync(S.m, 0, SUR_NPARAM * sizeof(float), ctx->stream));
  HIPCHK(hipMemsetAsync(S.v, 0, SUR_NPARAM * sizeof(float), ctx->stream));
  S.adam_t = 0;
  S.tab_valid = false;
  TRY(sur_refresh(ctx));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  S.have_params = true;
  return GD_OK;
}

gd_status gd_sur_tabulate(gd_ctx *ctx) {
  if (!ctx) return GD_ERR_INVALID;
  auto &S = ctx->sur;
  if (S.g.size() != ctx->cx.size()) return fail(ctx, GD_ERR_STATE, "call gd_sur_build_grids first");
  if (!S.have_params) return fail(ctx, GD_ERR_STATE, "no surrogate parameters");
  HIPCHK(hipSetDevice(ctx->device));
  SurW W{S.W1T, S.W2T, S.W2R, S.p + SP_B2, S.p + SP_W3, S.p + SP_B3};
  int *d_types = nullptr;
  HIPCHK(hipMalloc(&d_types, NLT * sizeof(int)));
  gd_status st = GD_OK;
  for (size_t c = 0; c < ctx->cx.size() && st == GD_OK; ++c) {
    HostComplex &h = ctx->cx[c];
    const int n = S.g[c].n;
    const size_t n3 = (size_t)n * n * n;
    if (!S.tab[c]) {
      if (hipMalloc(&S.tab[c], (size_t)h.d.n_prt * NLT * n3 * sizeof(float)) != hipSuccess) {
        S.tab[c] = nullptr;
        st = fail(ctx, GD_ERR_NOMEM, "surrogate table of complex %zu", c);
        break;
      }
    }
    // only the ligand's atom types are evaluated
    std::vector<int> types;
    for (int t = 0; t < NLT; ++t)
      for (int l = 0; l < h.d.lna; ++l)
        if (h.lig_t[l] == t) {
          types.push_back(t);
          break;
        }
    if (hipMemcpy(d_types, types.data(), types.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess) {
      st = fail(ctx, GD_ERR_HIP, "type list upload");
      break;
    }
    const int64_t jobs = (int64_t)((n3 + 15) / 16) * (int64_t)types.size();
    for (int p = 0; p < h.d.n_prt; ++p) {
      hipLaunchKernelGGL(gd_sur_tab_kernel, dim3((unsigned)((jobs + 3) / 4)), dim3(256), 0,
                         ctx->stream, S.grid[c] + (size_t)p * n3 * 32, (int)n3, d_types,
                         (int)types.size(), W, S.tab[c] + (size_t)p * NLT * n3);
    }
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(ctx->stream) != hipSuccess) {
      st = fail(ctx, GD_ERR_HIP, "tabulation kernel");
      break;
    }
    h.dev.stab = S.tab[c];
    ctx->cx_dirty = true;
  }
  (void)hipFree(d_types);
  TRY(st);
  TRY(sync_complexes(ctx));
  S.tab_valid = true;
  return GD_OK;
}

gd_status gd_sur_get_params(gd_ctx *ctx, float *params) {
  if (!ctx || !params) return GD_ERR_INVALID;
  if (!ctx->sur.have_params) return fail(ctx, GD_ERR_STATE, "no surrogate parameters");
  HIPCHK(hipSetDevice(ctx->device));
  HIPCHK(hipMemcpyAsync(params, ctx->sur.p, SUR_NPARAM * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return GD_OK;
}

gd_status gd_sur_dataset_clear(gd_ctx *ctx) {
  if (!ctx) return GD_ERR_INVALID;
  auto &S = ctx->sur;
  S.rows = 0;
  S.off.assign(1, 0);
  S.target.clear();
  S.pose_cx.clear();
  S.dirty = true;
  return GD_OK;
}

gd_status gd_sur_dataset_reserve(gd_ctx *ctx, int64_t rows) {
  if (!ctx || rows < 0) return GD_ERR_INVALID;
  auto &S = ctx->sur;
  if (rows <= S.cap) return GD_OK;
  HIPCHK(hipSetDevice(ctx->device));
  unsigned short *X = nullptr;
  const int64_t cap = (rows + 63) / 64 * 64;
  HIPCHK(hipMalloc(&X, (size_t)(cap + 64) * SUR_F * 2));
  HIPCHK(hipMemsetAsync(X, 0, (size_t)(cap + 64) * SUR_F * 2, ctx->stream));
  if (S.rows > 0)
    HIPCHK(hipMemcpyAsync(X, S.X, (size_t)S.rows * SUR_F * 2, hipMemcpyDeviceToDevice, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  if (S.X) (void)hipFree(S.X);
  S.X = X;
  S.cap = cap;
  S.dirty = true;
  return GD_OK;
}

gd_status gd_sur_dataset_add(gd_ctx *ctx, int cid, int n, const int *idx_lig, const int *idx_prt,
                             const float *movematrix, const float *target) {
  if (!ctx || cid < 0 || cid >= (int)ctx->cx.size() || n < 0) return GD_ERR_INVALID;
  if ((int)ctx->sur.g.size() <= cid) return fail(ctx, GD_ERR_STATE, "call gd_sur_build_grids first");
  if (n == 0) return GD_OK;
  if (!idx_lig || !idx_prt || !movematrix || !target) return GD_ERR_INVALID;
  const HostComplex &h = ctx->cx[cid];
  for (int i = 0; i < n; ++i)
    if (idx_lig[i] < 0 || idx_lig[i] >= h.d.n_lig || idx_prt[i] < 0 || idx_prt[i] >= h.d.n_prt)
      return fail(ctx, GD_ERR_INVALID, "pose %d: conformer index out of range", i);
  auto &S = ctx->sur;
  const int64_t add = (int64_t)n * h.d.lna;
  if (S.rows + add > S.cap) TRY(gd_sur_dataset_reserve(ctx, std::max<int64_t>(S.rows + add, 2 * S.cap)));
  HIPCHK(hipSetDevice(ctx->device));
  std::vector<void *> keep;
  int *dl, *dp;
  float *dm;
  gd_status s = GD_OK;
  if ((s = tmp_dev(ctx, keep, &dl, idx_lig, (size_t)n)) || (s = tmp_dev(ctx, keep, &dp, idx_prt, (size_t)n)) ||
      (s = tmp_dev(ctx, keep, &dm, movematrix, (size_t)n * 6))) {
    free_all(keep);
    return s;
  }
  hipLaunchKernelGGL(gd_sur_feat_kernel, dim3((unsigned)((add + 255) / 256)), dim3(256), 0,
                     ctx->stream, ctx->d_cx, cid, S.g[cid], n, dl, dp, dm, S.rows, S.X, nullptr,
                     S.cap + 64, (float *)nullptr);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  free_all(keep);
  if (e != hipSuccess) return fail(ctx, GD_ERR_HIP, "gd_sur_dataset_add: %s", hipGetErrorString(e));
  for (int i = 0; i < n; ++i) {
    S.off.push_back(S.off.back() + h.d.lna);
    S.target.push_back(target[i]);
    S.pose_cx.push_back(cid);
  }
  S.rows += add;
  S.dirty = true;
  return GD_OK;
}

gd_status gd_sur_dataset_add_multi(gd_ctx *ctx, int n, const int *cid, const int *idx_lig,
                                   const int *idx_prt, const float *movematrix,
                                   const float *target) {
  if (!ctx || n < 0) return GD_ERR_INVALID;
  if (n == 0) return GD_OK;
  if (!cid || !idx_lig || !idx_prt || !movematrix || !target) return GD_ERR_INVALID;
  auto &S = ctx->sur;
  if (S.g.size() != ctx->cx.size()) return fail(ctx, GD_ERR_STATE, "call gd_sur_build_grids first");
  std::vector<int64_t> row_of(n);
  int64_t add = 0;
  for (int i = 0; i < n; ++i) {
    if (cid[i] < 0 || cid[i] >= (int)ctx->cx.size())
      return fail(ctx, GD_ERR_INVALID, "pose %d: complex id out of range", i);
    const HostComplex &h = ctx->cx[cid[i]];
    if (idx_lig[i] < 0 || idx_lig[i] >= h.d.n_lig || idx_prt[i] < 0 || idx_prt[i] >= h.d.n_prt)
      return fail(ctx, GD_ERR_INVALID, "pose %d: conformer index out of range", i);
    row_of[i] = S.rows + add;
    add += h.d.lna;
  }
  if (S.rows + add > S.cap) TRY(gd_sur_dataset_reserve(ctx, std::max<int64_t>(S.rows + add, 2 * S.cap)));
  HIPCHK(hipSetDevice(ctx->device));
  TRY(sync_complexes(ctx));
  std::vector<void *> keep;
  int *dc, *dl, *dp;
  float *dm;
  int64_t *dr;
  gd_status s = GD_OK;
  if ((s = tmp_dev(ctx, keep, &dc, cid, (size_t)n)) || (s = tmp_dev(ctx, keep, &dl, idx_lig, (size_t)n)) ||
      (s = tmp_dev(ctx, keep, &dp, idx_prt, (size_t)n)) ||
      (s = tmp_dev(ctx, keep, &dm, movematrix, (size_t)n * 6)) ||
      (s = tmp_dev(ctx, keep, &dr, row_of.data(), (size_t)n))) {
    free_all(keep);
    return s;
  }
  hipLaunchKernelGGL(gd_sur_feat_multi_kernel, dim3((unsigned)n), dim3(64), 0, ctx->stream,
                     ctx->d_cx, n, dc, dl, dp, dm, dr, S.X, nullptr, S.cap + 64);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  free_all(keep);
  if (e != hipSuccess) return fail(ctx, GD_ERR_HIP, "gd_sur_dataset_add_multi: %s", hipGetErrorString(e));
  for (int i = 0; i < n; ++i) {
    S.off.push_back(S.off.back() + ctx->cx[cid[i]].d.lna);
    S.target.push_back(target[i]);
    S.pose_cx.push_back(cid[i]);
  }
  S.rows += add;
  S.dirty = true;
  return GD_OK;
}

int64_t gd_sur_dataset_size(const gd_ctx *ctx) { return ctx ? (int64_t)ctx->sur.target.size() : -1; }

gd_status gd_sur_loss_grad(gd_ctx *ctx, int pose_begin, int pose_end, double *loss, float *grad) {
  if (!ctx) return GD_ERR_INVALID;
  auto &S = ctx->sur;
  if (!S.have_params) return fail(ctx, GD_ERR_STATE, "no surrogate parameters");
  if (pose_begin < 0 || pose_end > (int)S.target.size() || pose_begin >= pose_end)
    return fail(ctx, GD_ERR_INVALID, "pose range [%d,%d) outside the dataset", pose_begin, pose_end);
  HIPCHK(hipSetDevice(ctx->device));
  TRY(sur_upload_index(ctx));
  TRY(sur_work(ctx, S.off[pose_end] - S.off[pose_begin], (int64_t)S.target.size()));
  TRY(sur_batch(ctx, pose_begin, pose_end, grad != nullptr, loss));
  if (grad) {
    HIPCHK(hipMemcpyAsync(grad, S.grad, SUR_NPARAM * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
  }
  return GD_OK;
}

gd_status gd_sur_predict(gd_ctx *ctx, int pose_begin, int pose_end, float *out) {
  if (!ctx || !out) return GD_ERR_INVALID;
  auto &S = ctx->sur;
  if (!S.have_params) return fail(ctx, GD_ERR_STATE, "no surrogate parameters");
  if (pose_begin < 0 || pose_end > (int)S.target.size() || pose_begin >= pose_end)
    return fail(ctx, GD_ERR_INVALID, "pose range outside the dataset");
  HIPCHK(hipSetDevice(ctx->device));
  TRY(sur_upload_index(ctx));
  TRY(sur_work(ctx, S.off[pose_end] - S.off[pose_begin], (int64_t)S.target.size()));
  // in chunks that bound the work buffers
  const int step = 65536;
  for (int p = pose_begin; p < pose_end; p += step)
    TRY(sur_batch(ctx, p, std::min(pose_end, p + step), false, nullptr));
  HIPCHK(hipMemcpyAsync(out, S.S + pose_begin, (size_t)(pose_end - pose_begin) * sizeof(float),
                        hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return GD_OK;
}

gd_status gd_sur_train(gd_ctx *ctx, int pose_begin, int pose_end, int epochs, int batch_poses,
                       float lr, double *epoch_loss, double *seconds) {
  if (!ctx || epochs < 0 || batch_poses <= 0 || !(lr > 0.0f)) return GD_ERR_INVALID;
  auto &S = ctx->sur;
  if (!S.have_params) return fail(ctx, GD_ERR_STATE, "no surrogate parameters");
  if (pose_begin < 0 || pose_end > (int)S.target.size() || pose_begin >= pose_end)
    return fail(ctx, GD_ERR_INVALID, "pose range outside the dataset");
  HIPCHK(hipSetDevice(ctx->device));
  TRY(sur_upload_index(ctx));
  int64_t maxrows = 0;
  for (int p = pose_begin; p < pose_end; p += batch_poses) {
    const int q = std::min(pose_end, p + batch_poses);
    maxrows = std::max<int64_t>(maxrows, S.off[q] - S.off[p]);
  }
  TRY(sur_work(ctx, maxrows, (int64_t)S.target.size()));
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const int n_batch = (pose_end - pose_begin + batch_poses - 1) / batch_poses;
  // per-batch sums of squared residuals (before each batch's update)
  float *d_bl = nullptr;
  HIPCHK(hipMalloc(&d_bl, (size_t)n_batch * sizeof(float)));
  std::vector<float> bl(n_batch);
  const auto t0 = std::chrono::steady_clock::now();
  S.tab_valid = false;  // the parameters change
  gd_status st = GD_OK;
  for (int ep = 0; ep < epochs && st == GD_OK; ++ep) {
    int b = 0;
    for (int p = pose_begin; p < pose_end && st == GD_OK; p += batch_poses, ++b) {
      const int q = std::min(pose_end, p + batch_poses);
      st = sur_batch(ctx, p, q, true, nullptr, true);
      if (st != GD_OK) break;
      hipLaunchKernelGGL(gd_mlp_sqsum_kernel, dim3(1), dim3(256), 0, ctx->stream, S.sq + p,
                         q - p, d_bl + b);
      S.adam_t += 1;
      const float c1 = 1.0f - std::pow(b1, (float)S.adam_t);
      const float c2 = 1.0f - std::pow(b2, (float)S.adam_t);
      hipLaunchKernelGGL(gd_mlp_update_kernel, dim3((SUR_NPARAM + 255) / 256), dim3(256), 0,
                         ctx->stream, S.p1, S.p2, S.pb2, S.pw3, S.last_groups, S.p, S.m, S.v,
                         S.grad, lr, b1, b2, eps, c1, c2, S.W1T, S.W2T, S.W2R);
      if (hipGetLastError() != hipSuccess) st = fail(ctx, GD_ERR_HIP, "training launch failed");
    }
    if (st != GD_OK || !epoch_loss) continue;
    if (ep + 1 < epochs) {
      // the epoch's batch losses, each taken before its batch's update
      if (hipMemcpyAsync(bl.data(), d_bl, (size_t)n_batch * sizeof(float),
                         hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
          hipStreamSynchronize(ctx->stream) != hipSuccess) {
        st = fail(ctx, GD_ERR_HIP, "batch loss copy failed");
        break;
      }
      double tot = 0.0;
      for (float v : bl) tot += v;
      epoch_loss[ep] = tot / (double)(pose_end - pose_begin);
    } else {
      // last epoch: the loss over the whole range with the final parameters
      double tot = 0.0;
      for (int p = pose_begin; p < pose_end && st == GD_OK; p += batch_poses) {
        const int q = std::min(pose_end, p + batch_poses);
        double l = 0.0;
        st = sur_batch(ctx, p, q, false, &l);
        tot += l * (q - p);
      }
      epoch_loss[ep] = tot / (double)(pose_end - pose_begin);
    }
  }
  (void)hipStreamSynchronize(ctx->stream);
  (void)hipFree(d_bl);
  TRY(st);
  HIPCHK(hipStreamSynchronize(ctx->stream));
  if (seconds) *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return GD_OK;
}

}  // extern "C"

#include "gd_post.hip"
