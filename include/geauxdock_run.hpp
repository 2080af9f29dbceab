// geauxdock_run.hpp -- drop-in C++ Run() for GeauxDock's host code.
//
// The reference's entry point (src/run.h:10-22, src/run.cu:33-289)
//
//   void Run(const Ligand*, const Protein*, const Psp*, const Kde*, const Mcs*,
//            const EnePara*, const Temp*, const Replica*, const McPara*, McLog*,
//            map<int, vector<LigRecordSingleStep>>&, const ComplexSize);
//
// becomes a call of geauxdock::RunT with the same arguments. RunT reads the
// reference's structures by field name (src/dock.h, src/dock_soa.h), so it
// compiles against the reference's own headers: a maintainer replaces the body
// of Run() in src/run.cu with `geauxdock::RunT(lig, prt, ..., complexsize);`
// and links libgeauxdock.so. Nothing here depends on CUDA.
//
// Behaviour follows Run() and its single-GPU launcher
// (src/kernel_cuda_lancher_siglegpu.cu:5-69): the start poses are evaluated
// and recorded at step 0 (MonteCarlo_Init_d), then MC steps run in windows of
// steps_per_dump, each window's accepted-state records are appended to
// multi_reps_records[replica], and McLog receives t0 (MC compute seconds), t1
// (wall seconds), steps_total and ar = records / (n_rep * steps_total).
// Differences, each deliberate: the run length is mcpara->steps_total steps
// (the reference loops until a record count is reached), the random stream is
// Philox keyed by RunOptions::seed, and temperature exchange runs only when
// RunOptions::exchange is set (the reference compiles it out, IS_EXCHANGE 0).
#pragma once

#include <chrono>
#include <cstdio>
#include <ctime>
#include <iomanip>
#include <fstream>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "geauxdock.h"
#include "geauxdock_io.h"

namespace geauxdock {

struct RunOptions {
  uint64_t seed = (uint64_t)std::time(nullptr);  // the reference seeds from time()
  bool exchange = false;                         // IS_EXCHANGE (src/toggle.h:12)
  int device = 0;
  int move_kind = GD_MOVE_UNIFORM;  // scale * (2U - 1), src/kernel_cuda_l1_montecarlo.cu:100
};

class Error : public std::runtime_error {
 public:
  Error(gd_status s, const std::string &m) : std::runtime_error(m), status(s) {}
  gd_status status;
};

namespace detail {
template <class A>
constexpr int extent_of(const A &) {
  return (int)(sizeof(A) / sizeof(A{}[0]));
}
inline void check(gd_ctx *ctx, gd_status s, const char *what) {
  if (s != GD_OK)
    throw Error(s, std::string(what) + ": " + (ctx ? gd_last_error(ctx) : "no context"));
}
template <class Record>
struct Sink {
  std::map<int, std::vector<Record>> *out;
  int64_t count = 0;
  static void call(void *user, int, const gd_record *r, int64_t n) {
    Sink *s = static_cast<Sink *>(user);
    for (int64_t i = 0; i < n; ++i) {
      Record rec;
      rec.replica.idx_rep = r[i].replica.idx_rep;
      rec.replica.idx_prt = r[i].replica.idx_prt;
      rec.replica.idx_tmp = r[i].replica.idx_tmp;
      rec.replica.idx_lig = r[i].replica.idx_lig;
      for (int k = 0; k < GD_MAXWEI; ++k) rec.energy.e[k] = r[i].energy.e[k];
      rec.energy.cms = r[i].energy.cms;
      rec.energy.rmsd = r[i].energy.rmsd;
      for (int k = 0; k < 6; ++k) rec.movematrix[k] = r[i].movematrix[k];
      rec.step = r[i].step;
      (*s->out)[r[i].replica.idx_rep].push_back(rec);
    }
    s->count += n;
  }
};
}  // namespace detail

// Copies the reference's structures into a complex description, runs it and
// fills McLog and the records map. Throws geauxdock::Error on failure.
template <class Ligand, class Protein, class Psp, class Kde, class Mcs, class EnePara,
          class Temp, class Replica, class McPara, class McLog, class Record, class ComplexSize>
void RunT(const Ligand *lig, const Protein *prt, const Psp *psp, const Kde *kde, const Mcs *mcs,
          const EnePara *enepara, const Temp *temp, const Replica *replica, const McPara *mcpara,
          McLog *mclog, std::map<int, std::vector<Record>> &multi_reps_records,
          const ComplexSize cs, const RunOptions &opt = RunOptions()) {
  (void)replica;  // the flat layout of SetReplica (src/util.C:842-861) is implied
  const auto t_wall0 = std::chrono::steady_clock::now();
  const int lna = cs.lna, pnp = cs.pnp, pnk = cs.pnk, pos = cs.pos;
  const int n_lig = cs.n_lig, n_prt = cs.n_prt, n_tmp = cs.n_tmp;

  // ligand conformers: lig[l] for l < n_lig is conformer l at p = 0, t = 0
  std::vector<float> lx((size_t)n_lig * lna), ly(lx.size()), lz(lx.size());
  for (int l = 0; l < n_lig; ++l)
    for (int a = 0; a < lna; ++a) {
      lx[(size_t)l * lna + a] = lig[l].coord_orig.x[a];
      ly[(size_t)l * lna + a] = lig[l].coord_orig.y[a];
      lz[(size_t)l * lna + a] = lig[l].coord_orig.z[a];
    }
  std::vector<int> lt(lna), ln(lna);
  std::vector<float> lc(lna);
  for (int a = 0; a < lna; ++a) {
    lt[a] = lig[0].t[a];
    lc[a] = lig[0].c[a];
    ln[a] = lig[0].n[a];
  }
  // protein conformers
  std::vector<float> px((size_t)n_prt * pnp), py(px.size()), pz(px.size());
  for (int p = 0; p < n_prt; ++p)
    for (int i = 0; i < pnp; ++i) {
      px[(size_t)p * pnp + i] = prt[p].x[i];
      py[(size_t)p * pnp + i] = prt[p].y[i];
      pz[(size_t)p * pnp + i] = prt[p].z[i];
    }
  std::vector<int> pt(pnp), pc(pnp), pseq(pnp), pflag(pnp);
  std::vector<float> pele(pnp), phpp(pnp);
  for (int i = 0; i < pnp; ++i) {
    pt[i] = prt[0].t[i];
    pc[i] = prt[0].c[i];
    pele[i] = prt[0].ele[i];
    pseq[i] = prt[0].seq3r[i];
    pflag[i] = prt[0].c0_and_d12_or_c2[i];
    phpp[i] = prt[0].hpp[i];
  }
  // MCS templates, structure of arrays
  const int n_mcs = pos > 0 ? detail::extent_of(mcs[0].x) : 1;
  std::vector<float> mx((size_t)pos * n_mcs), my(mx.size()), mz(mx.size()), tcc(pos);
  for (int m = 0; m < pos; ++m) {
    tcc[m] = mcs[m].tcc;
    for (int j = 0; j < n_mcs; ++j) {
      mx[(size_t)m * n_mcs + j] = mcs[m].x[j];
      my[(size_t)m * n_mcs + j] = mcs[m].y[j];
      mz[(size_t)m * n_mcs + j] = mcs[m].z[j];
    }
  }
  // energy parameters, field by field
  gd_enepara ep;
  for (int i = 0; i < GD_MAXTP2; ++i)
    for (int j = 0; j < GD_MAXTP1; ++j) {
      ep.p1a[i][j] = enepara->p1a[i][j];
      ep.p2a[i][j] = enepara->p2a[i][j];
      ep.pmf0[i][j] = enepara->pmf0[i][j];
      ep.pmf1[i][j] = enepara->pmf1[i][j];
      ep.hdb0[i][j] = enepara->hdb0[i][j];
      ep.hdb1[i][j] = enepara->hdb1[i][j];
    }
  ep.lj0 = enepara->lj0;
  ep.lj1 = enepara->lj1;
  ep.el0 = enepara->el0;
  ep.el1 = enepara->el1;
  for (int i = 0; i < GD_MAXTP3; ++i) ep.ele[i] = enepara->ele[i];
  ep.a1 = enepara->a1;
  ep.b1 = enepara->b1;
  for (int i = 0; i < GD_MAXTP4; ++i) ep.hpp[i] = enepara->hpp[i];
  for (int i = 0; i < GD_MAXTP2; ++i) {
    ep.hpl0[i] = enepara->hpl0[i];
    ep.hpl1[i] = enepara->hpl1[i];
    ep.hpl2[i] = enepara->hpl2[i];
  }
  ep.kde2 = enepara->kde2;
  ep.kde3 = enepara->kde3;
  for (int i = 0; i < GD_MAXWEI; ++i) {
    ep.w[i] = enepara->w[i];
    ep.a_para[i] = enepara->a_para[i];
    ep.b_para[i] = enepara->b_para[i];
  }
  std::vector<gd_temp> tmp(n_tmp);
  for (int t = 0; t < n_tmp; ++t) {
    tmp[t].t = temp[t].t;
    tmp[t].minus_beta = temp[t].minus_beta;
    tmp[t].order = temp[t].order;
  }

  gd_complex_desc d{};
  d.n_lig = n_lig;
  d.n_prt = n_prt;
  d.n_tmp = n_tmp;
  d.lna = lna;
  d.pnp = pnp;
  d.pnk = pnk;
  d.pos = pos;
  d.n_res = detail::extent_of(psp->psp[0]);
  d.n_mcs = n_mcs;
  d.lig_x = lx.data();
  d.lig_y = ly.data();
  d.lig_z = lz.data();
  d.lig_t = lt.data();
  d.lig_c = lc.data();
  d.lig_n = ln.data();
  for (int i = 0; i < 3; ++i) d.pocket_center[i] = lig[0].coord_orig.center[i];
  d.prt_x = px.data();
  d.prt_y = py.data();
  d.prt_z = pz.data();
  d.prt_t = pt.data();
  d.prt_c = pc.data();
  d.prt_ele = pele.data();
  d.prt_seq3r = pseq.data();
  d.prt_c0_and_d12_or_c2 = pflag.data();
  d.prt_hpp = phpp.data();
  d.psp = &psp->psp[0][0];
  d.kde_x = kde->x;
  d.kde_y = kde->y;
  d.kde_z = kde->z;
  d.kde_t = kde->t;
  d.mcs_x = mx.data();
  d.mcs_y = my.data();
  d.mcs_z = mz.data();
  d.mcs_tcc = tcc.data();
  d.enepara = &ep;
  d.temp = tmp.data();
  for (int i = 0; i < 6; ++i) d.move_scale[i] = mcpara->move_scale[i];
  d.move_kind = opt.move_kind;

  gd_opts o{};
  o.device = opt.device;
  o.seed = opt.seed;
  o.steps_per_dump = mcpara->steps_per_dump;
  o.record = 1;
  gd_ctx *ctx = nullptr;
  gd_status s = gd_create(&ctx, &o);
  if (s != GD_OK) {
    const std::string m = ctx ? gd_last_error(ctx) : "gd_create";
    gd_destroy(ctx);
    throw Error(s, m);
  }
  try {
    int cid = -1;
    detail::check(ctx, gd_add_complex(ctx, &d, &cid), "gd_add_complex");
    detail::Sink<Record> sink{&multi_reps_records};
    gd_run_stats st{};
    detail::check(ctx,
                  gd_run(ctx, mcpara->steps_total, opt.exchange ? mcpara->steps_per_exchange : 0,
                         &detail::Sink<Record>::call, &sink, &st),
                  "gd_run");
    mclog->t0 += st.t_compute;
    mclog->steps_total = mcpara->steps_total;
    const double trials = (double)cs.n_rep * (double)mcpara->steps_total;
    mclog->ar = trials > 0 ? (float)((double)sink.count / trials) : 0.0f;
  } catch (...) {
    gd_destroy(ctx);
    throw;
  }
  gd_destroy(ctx);
  mclog->t1 += std::chrono::duration<double>(std::chrono::steady_clock::now() - t_wall0).count();
}

// The performance part of PrintSummary (src/util.C:1281-1364): sizes, the
// acceptance ratio and replica-steps per second over compute time (t0) and
// wall time (t1), with the reference's "speedup over 843.75" lines.
template <class McLog, class ComplexSize>
void PrintSummary(const McLog *mclog, const ComplexSize &cs, FILE *out = stdout) {
  const char *bar =
      "================================================================================\n";
  std::fprintf(out, "%s", bar);
  std::fprintf(out, "lig conformations\t\t%d\n", cs.n_lig);
  std::fprintf(out, "prt conformations\t\t%d\n", cs.n_prt);
  std::fprintf(out, "temperatures\t\t\t%d\n", cs.n_tmp);
  std::fprintf(out, "replica ensembles\t\t%d\n", cs.n_rep);
  std::fprintf(out, "size_lig\t\t\t%d\nsize_prt\t\t\t%d\nsize_pnk\t\t\t%d\nsize_mcs\t\t\t%d\n",
               cs.lna, cs.pnp, cs.pnk, cs.pos);
  std::fprintf(out, "AR of MC\t\t\t%.3f\n%sPerformance\n%s", mclog->ar, bar, bar);
  const double n = (double)mclog->steps_total * cs.n_rep;
  const double times[2] = {mclog->t0, mclog->t1};
  const char *names[2] = {"compute time", "wall time"};
  for (int i = 0; i < 2; ++i) {
    const double rate = times[i] > 0 ? n / times[i] : 0.0;
    std::fprintf(out, "%s\t\t\t%.3f seconds\n", names[i], times[i]);
    std::fprintf(out, "time per MC sweep per replica\t%.3f * 1e-6 seconds\n",
                 rate > 0 ? 1e6 / rate : 0.0);
    std::fprintf(out, "MC sweeps per second\t\t%.3f\n", rate);
    std::fprintf(out, "speedup over 843.75\t\t%.3f X\n", rate / 843.75);
  }
  std::fprintf(out, "%s", bar);
}

// printStates (src/util.C:1504-1550) through gd_write_states_csv: one line per
// record, "lig prt cms rmsd mv1..mv6 vdw ele pmf psp hdb hpc kde lhm edst ener",
// 4 decimals.
template <class Record>
void WriteStates(const std::vector<Record> &steps, const std::string &path) {
  std::vector<gd_record> recs(steps.size());
  for (size_t i = 0; i < steps.size(); ++i) {
    const Record &s = steps[i];
    gd_record &r = recs[i];
    r.replica.idx_rep = s.replica.idx_rep;
    r.replica.idx_prt = s.replica.idx_prt;
    r.replica.idx_tmp = s.replica.idx_tmp;
    r.replica.idx_lig = s.replica.idx_lig;
    for (int k = 0; k < GD_MAXWEI; ++k) r.energy.e[k] = s.energy.e[k];
    r.energy.cms = s.energy.cms;
    r.energy.rmsd = s.energy.rmsd;
    for (int k = 0; k < 6; ++k) r.movematrix[k] = s.movematrix[k];
    r.step = s.step;
  }
  char err[512];
  const gd_status st = gd_write_states_csv(path.c_str(), recs.data(), (int64_t)recs.size(), err,
                                           sizeof err);
  if (st != GD_OK) throw Error(st, err);
}

}  // namespace geauxdock
