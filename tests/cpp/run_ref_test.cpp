// Exercises include/geauxdock_run.hpp the way GeauxDock's main() would
// (src/dock.C:150-192): structures with the reference's field names and
// fixed capacities, filled as OptimizeLigand / InitLigCoord / SetReplica
// leave them, then Run(). Input and output are raw binary files written and
// read by tests/test_gpu_run_ref.py.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <vector>

#include "geauxdock_run.hpp"

// Capacities of the reference build (src/size.h:19-122).
constexpr int MAXPRO = 10000, MAXLIG = 100, MAXKDE = 10000, MAXMCS = 256, MAXREP = 2048;
constexpr int MAXTP1 = 30, MAXTP2 = 24, MAXTP3 = 50, MAXTP4 = 20, MAXWEI = 10;

struct LigCoord {
  float x[MAXLIG], y[MAXLIG], z[MAXLIG];
  float center[3];
};
struct Energy {
  float e[MAXWEI];
  float cms, rmsd;
};
struct Ligand {
  LigCoord coord_orig, coord_new;
  float movematrix_old[6], movematrix_new[6];
  Energy energy_old, energy_new;
  int is_move_accepted;
  int t[MAXLIG];
  float c[MAXLIG];
  int n[MAXLIG];
  int lna;
};
struct Protein {
  float x[MAXPRO], y[MAXPRO], z[MAXPRO];
  int t[MAXPRO], c[MAXPRO];
  float ele[MAXPRO];
  int seq3r[MAXPRO], c0_and_d12_or_c2[MAXPRO];
  float hpp[MAXPRO];
  int pnp;
  float pocket_center[3];
};
struct Psp {
  float psp[MAXLIG][MAXPRO];
};
struct Kde {
  float x[MAXKDE], y[MAXKDE], z[MAXKDE];
  int t[MAXKDE];
  int pnk;
};
struct Mcs {
  float x[MAXMCS], y[MAXMCS], z[MAXMCS];
  float tcc;
};
struct EnePara {
  float p1a[MAXTP2][MAXTP1], p2a[MAXTP2][MAXTP1];
  float lj0, lj1, el0, el1, ele[MAXTP3], a1, b1;
  float pmf0[MAXTP2][MAXTP1], pmf1[MAXTP2][MAXTP1], hdb0[MAXTP2][MAXTP1], hdb1[MAXTP2][MAXTP1];
  float hpp[MAXTP4], hpl0[MAXTP2], hpl1[MAXTP2], hpl2[MAXTP2];
  float kde2, kde3;
  float w[MAXWEI], a_para[MAXWEI], b_para[MAXWEI];
};
struct Temp {
  float t, minus_beta;
  int order;
};
struct Replica {
  int idx_rep, idx_prt, idx_tmp, idx_lig;
};
struct McPara {
  int steps_total, steps_per_dump, steps_per_exchange;
  float move_scale[6];
  char hdf_path[512], csv_path[512];
};
struct McLog {
  double t0, t1, t2;
  int ac_temp_exchg;
  int acs_temp_exchg[MAXREP];
  int ac_mc;
  float ar;
  int steps_total;
};
struct LigRecordSingleStep {
  Replica replica;
  Energy energy;
  float movematrix[6];
  int step;
};
struct ComplexSize {
  int n_lig, n_prt, n_tmp, n_rep, lna, pnp, pnk, pos;
};

template <class T>
static void rd(FILE *f, T *p, size_t n) {
  if (fread(p, sizeof(T), n, f) != n) {
    fprintf(stderr, "short read\n");
    exit(3);
  }
}

int main(int argc, char **argv) {
  if (argc < 4) {
    fprintf(stderr, "usage: %s in.bin out.bin states.csv\n", argv[0]);
    return 2;
  }
  FILE *f = fopen(argv[1], "rb");
  if (!f) return 2;
  int32_t h[16];
  rd(f, h, 16);
  const int n_lig = h[0], n_prt = h[1], n_tmp = h[2], lna = h[3], pnp = h[4], pnk = h[5];
  const int pos = h[6], n_res = h[7], n_mcs = h[8];
  ComplexSize cs{n_lig, n_prt, n_tmp, n_lig * n_prt * n_tmp, lna, pnp, pnk, pos};
  McPara mc{};
  mc.steps_total = h[9];
  mc.steps_per_dump = h[10];
  mc.steps_per_exchange = h[11];
  geauxdock::RunOptions opt;
  opt.exchange = h[12] != 0;
  opt.seed = (uint64_t)(uint32_t)h[13] | ((uint64_t)(uint32_t)h[14] << 32);
  opt.move_kind = h[15];
  rd(f, mc.move_scale, 6);
  float pocket[3];
  rd(f, pocket, 3);

  std::vector<Ligand> lig(cs.n_rep);
  std::vector<float> buf((size_t)n_lig * lna);
  for (int axis = 0; axis < 3; ++axis) {
    rd(f, buf.data(), buf.size());
    for (int l = 0; l < n_lig; ++l)
      for (int a = 0; a < lna; ++a) {
        float *dst = axis == 0 ? lig[l].coord_orig.x : axis == 1 ? lig[l].coord_orig.y : lig[l].coord_orig.z;
        dst[a] = buf[(size_t)l * lna + a];
      }
  }
  std::vector<int> it(lna);
  std::vector<float> fc(lna);
  rd(f, it.data(), lna);
  for (int l = 0; l < n_lig; ++l) memcpy(lig[l].t, it.data(), lna * sizeof(int));
  rd(f, fc.data(), lna);
  for (int l = 0; l < n_lig; ++l) memcpy(lig[l].c, fc.data(), lna * sizeof(float));
  rd(f, it.data(), lna);
  for (int l = 0; l < n_lig; ++l) {
    memcpy(lig[l].n, it.data(), lna * sizeof(int));
    lig[l].lna = lna;
    for (int i = 0; i < 3; ++i) lig[l].coord_orig.center[i] = pocket[i];  // OptimizeLigand
    for (int i = 0; i < 6; ++i) lig[l].movematrix_old[i] = 0.0f;          // InitLigCoord
  }
  for (int p = 0; p < n_prt; ++p)  // SetReplica duplicates ligand conformers
    for (int t = 0; t < n_tmp; ++t)
      for (int l = 0; l < n_lig; ++l) lig[n_tmp * n_lig * p + n_lig * t + l] = lig[l];

  std::unique_ptr<Protein[]> prt(new Protein[n_prt]);
  std::vector<float> pb((size_t)n_prt * pnp);
  for (int axis = 0; axis < 3; ++axis) {
    rd(f, pb.data(), pb.size());
    for (int p = 0; p < n_prt; ++p) {
      float *dst = axis == 0 ? prt[p].x : axis == 1 ? prt[p].y : prt[p].z;
      memcpy(dst, pb.data() + (size_t)p * pnp, pnp * sizeof(float));
    }
  }
  std::vector<int> ti(pnp), ci(pnp), si(pnp), fi(pnp);
  std::vector<float> ef(pnp), hf(pnp);
  rd(f, ti.data(), pnp);
  rd(f, ci.data(), pnp);
  rd(f, ef.data(), pnp);
  rd(f, si.data(), pnp);
  rd(f, fi.data(), pnp);
  rd(f, hf.data(), pnp);
  for (int p = 0; p < n_prt; ++p) {
    memcpy(prt[p].t, ti.data(), pnp * 4);
    memcpy(prt[p].c, ci.data(), pnp * 4);
    memcpy(prt[p].ele, ef.data(), pnp * 4);
    memcpy(prt[p].seq3r, si.data(), pnp * 4);
    memcpy(prt[p].c0_and_d12_or_c2, fi.data(), pnp * 4);
    memcpy(prt[p].hpp, hf.data(), pnp * 4);
    prt[p].pnp = pnp;
    for (int i = 0; i < 3; ++i) prt[p].pocket_center[i] = pocket[i];
  }
  std::unique_ptr<Psp> psp(new Psp());
  std::vector<float> ps((size_t)MAXTP2 * n_res);
  rd(f, ps.data(), ps.size());
  for (int lt = 0; lt < MAXTP2; ++lt)
    memcpy(psp->psp[lt], ps.data() + (size_t)lt * n_res, n_res * sizeof(float));
  std::unique_ptr<Kde> kde(new Kde());
  rd(f, kde->x, pnk);
  rd(f, kde->y, pnk);
  rd(f, kde->z, pnk);
  rd(f, kde->t, pnk);
  kde->pnk = pnk;
  std::vector<Mcs> mcs(pos > 0 ? pos : 1);
  for (auto &m : mcs)
    for (int j = 0; j < MAXMCS; ++j) m.x[j] = m.y[j] = m.z[j] = GD_MCS_INVALID_COORD;
  std::vector<float> mb((size_t)pos * n_mcs);
  for (int axis = 0; axis < 3; ++axis) {
    rd(f, mb.data(), mb.size());
    for (int m = 0; m < pos; ++m) {
      float *dst = axis == 0 ? mcs[m].x : axis == 1 ? mcs[m].y : mcs[m].z;
      memcpy(dst, mb.data() + (size_t)m * n_mcs, n_mcs * sizeof(float));
    }
  }
  std::vector<float> tcc(pos);
  rd(f, tcc.data(), pos);
  for (int m = 0; m < pos; ++m) mcs[m].tcc = tcc[m];
  EnePara ep;
  rd(f, reinterpret_cast<char *>(&ep), sizeof(EnePara));
  std::vector<Temp> temp(n_tmp);
  for (auto &t : temp) rd(f, reinterpret_cast<char *>(&t), sizeof(Temp));
  fclose(f);
  std::vector<Replica> rep(cs.n_rep);
  for (int p = 0; p < n_prt; ++p)
    for (int t = 0; t < n_tmp; ++t)
      for (int l = 0; l < n_lig; ++l) rep[n_tmp * n_lig * p + n_lig * t + l] = {n_tmp * n_lig * p + n_lig * t + l, p, t, l};

  std::unique_ptr<McLog> log(new McLog());
  std::map<int, std::vector<LigRecordSingleStep>> records;
  try {
    geauxdock::RunT(lig.data(), prt.get(), psp.get(), kde.get(), mcs.data(), &ep, temp.data(),
                    rep.data(), &mc, log.get(), records, cs, opt);
  } catch (const geauxdock::Error &e) {
    fprintf(stderr, "Run failed: %s\n", e.what());
    return 1;
  }
  FILE *o = fopen(argv[2], "wb");
  if (!o) return 2;
  int64_t n = 0;
  for (auto &kv : records) n += (int64_t)kv.second.size();
  fwrite(&n, sizeof n, 1, o);
  for (auto &kv : records) fwrite(kv.second.data(), sizeof(LigRecordSingleStep), kv.second.size(), o);
  const double lg[3] = {log->t0, log->t1, (double)log->ar};
  fwrite(lg, sizeof lg, 1, o);
  fclose(o);
  if (!records.empty()) geauxdock::WriteStates(records.begin()->second, argv[3]);
  geauxdock::PrintSummary(log.get(), cs, stderr);
  printf("records %lld ar %.4f t0 %.4f t1 %.4f steps %d\n", (long long)n, log->ar, log->t0, log->t1,
         log->steps_total);
  return 0;
}
