// gd_io.cpp -- GeauxDock input formats into a gd_complex_desc (host C++).
//
// Formats and derivations follow the reference's loaders so that the same
// input files give the same complex; the type tables below are the
// reference's code numbering (src/data.C), kept identical because the
// energy parameter tables are indexed by them.
#include "../../include/geauxdock_io.h"

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <memory>
#include <new>
#include <stdexcept>
#include <sstream>
#include <string>
#include <vector>

namespace {

// ligand atom types, src/data.C:532-583 (index = code)
const char *const kLigTypes[GD_MAXTP2] = {"Br",   "C.1",  "C.2",   "C.3",  "C.ar", "C.cat",
                                          "Cl",   "F",    "I",     "N.1",  "N.2",  "N.3",
                                          "N.4",  "N.am", "N.ar",  "N.pl3", "O.2", "O.3",
                                          "O.co2", "P.3", "S.2",   "S.3",  "S.O",  "S.O2"};
// residues, src/data.C:239-287 (three-letter) and :290-337 (one-letter)
const char *const kRes3[GD_MAXTP4] = {"ALA", "ARG", "CYS", "PRO", "LEU", "ILE", "GLU",
                                      "HIS", "TRP", "THR", "MET", "ASP", "GLY", "PHE",
                                      "LYS", "GLN", "SER", "ASN", "TYR", "VAL"};
const char kRes1[GD_MAXTP4 + 1] = "ARCPLIEHWTMDGFKQSNYV";
// coarse receptor point types, src/data.C:391-459
const char *const kPntTypes[GD_MAXTP1] = {"CA", "PP", "A1", "C1", "D1", "E1", "E2", "F1",
                                          "F2", "H1", "H2", "I1", "K1", "K2", "L1", "M1",
                                          "M2", "N1", "P1", "Q1", "Q2", "R1", "R2", "S1",
                                          "T1", "V1", "W1", "W2", "Y1", "Y2"};

// Side-chain points of each residue: (point type, class), in the order the
// reference creates them (src/load.C:980-1051). GLY has none.
struct SidePts {
  int n;
  int t[2], c[2];
};
const SidePts kSide[GD_MAXTP4] = {
    {1, {2, 0}, {2, 0}},   {2, {21, 22}, {2, 3}}, {1, {3, 0}, {2, 0}},   {1, {18, 0}, {2, 0}},
    {1, {14, 0}, {2, 0}},  {1, {11, 0}, {2, 0}},  {2, {5, 6}, {2, 3}},   {2, {9, 10}, {2, 3}},
    {2, {26, 27}, {2, 3}}, {1, {24, 0}, {2, 0}},  {2, {15, 16}, {2, 3}}, {1, {4, 0}, {2, 0}},
    {0, {0, 0}, {0, 0}},   {2, {7, 8}, {2, 3}},   {2, {12, 13}, {2, 3}}, {2, {19, 20}, {2, 3}},
    {1, {23, 0}, {2, 0}},  {1, {17, 0}, {2, 0}},  {2, {28, 29}, {2, 3}}, {1, {25, 0}, {2, 0}}};

// Which atoms a point averages (src/load.C:1093-1163):
//   CA      the residue's CA;
//   SIDE    a one-centroid side chain: every atom but N, CA, C, O;
//   DISTAL  the far centroid of a two-point side chain: also without CB, CG;
//   PROX    the near centroid of a two-point side chain: CB and CG;
//   PEPTIDE N of the residue and C, O of the previous one.
enum PointRule { R_CA, R_SIDE, R_DISTAL, R_PROX, R_PEPTIDE, R_NONE };
PointRule rule_of(int t) {
  switch (t) {
    case 0: return R_CA;
    case 1: return R_PEPTIDE;
    case 2: case 3: case 4: case 11: case 14: case 17: case 18: case 23: case 24: case 25:
      return R_SIDE;
    case 5: case 7: case 9: case 12: case 15: case 19: case 21: case 26: case 28:
      return R_DISTAL;
    case 6: case 8: case 10: case 13: case 16: case 20: case 22: case 27: case 29:
      return R_PROX;
    default: return R_NONE;
  }
}

int lig_code(const std::string &s) {
  for (int i = 0; i < GD_MAXTP2; ++i)
    if (s == kLigTypes[i]) return i;
  return -1;
}
int res3_code(const std::string &s) {
  for (int i = 0; i < GD_MAXTP4; ++i)
    if (s == kRes3[i]) return i;
  return -1;
}
int res1_code(char c) {
  const char *p = std::strchr(kRes1, c);
  return (p && c) ? (int)(p - kRes1) : -1;
}
int pnt_code(const std::string &s) {
  for (int i = 0; i < GD_MAXTP1; ++i)
    if (s == kPntTypes[i]) return i;
  return -1;
}

std::vector<std::string> tokens(const std::string &line) {
  std::vector<std::string> out;
  std::istringstream ss(line);
  std::string t;
  while (ss >> t) out.push_back(t);
  return out;
}

bool starts(const std::string &s, const char *p) { return s.compare(0, std::strlen(p), p) == 0; }

struct Err {
  char *buf;
  size_t len;
  gd_status fail(gd_status st, const char *fmt, ...) {
    if (buf && len) {
      va_list ap;
      va_start(ap, fmt);
      std::vsnprintf(buf, len, fmt, ap);
      va_end(ap);
    }
    return st;
  }
};

bool read_lines(const char *path, std::vector<std::string> &lines) {
  std::ifstream f(path);
  if (!f.is_open()) return false;
  std::string l;
  while (std::getline(f, l)) {
    if (!l.empty() && l.back() == '\r') l.pop_back();
    lines.push_back(l);
  }
  return true;
}

float fnum(const std::string &s) { return (float)std::atof(s.c_str()); }
float fcol(const std::string &line, size_t pos, size_t n) {
  return pos < line.size() ? (float)std::atof(line.substr(pos, n).c_str()) : 0.0f;
}
int icol(const std::string &line, size_t pos, size_t n) {
  return pos < line.size() ? std::atoi(line.substr(pos, n).c_str()) : 0;
}

// ---- energy parameters (EnePara0 of src/dock.h:256-272) -------------------
struct RawPara {
  float vdw[GD_MAXTP1][GD_MAXTP2][2] = {};
  float pmf[GD_MAXTP1][GD_MAXTP2][2] = {};
  float hdb[GD_MAXTP1][GD_MAXTP2][2] = {};
  float lj[3] = {};
  float el[2] = {};
  float ele[GD_MAXTP3] = {};
  float hpp[GD_MAXTP4] = {};
  float hpl[GD_MAXTP2][2] = {};
  float kde = 0.0f;
  float w[GD_MAXWEI] = {}, a[GD_MAXWEI] = {}, b[GD_MAXWEI] = {};
};

gd_status load_paras(const char *path, RawPara &p, Err &e) {
  std::vector<std::string> lines;
  if (!read_lines(path, lines)) return e.fail(GD_ERR_INVALID, "cannot open %s", path);
  for (const std::string &l : lines) {
    if (l.size() <= 3) continue;
    const std::vector<std::string> t = tokens(l);
    if (t.empty()) continue;
    const std::string &k = t[0];
    auto pair_table = [&](float (*tab)[GD_MAXTP2][2]) -> gd_status {
      if (t.size() < 5) return e.fail(GD_ERR_INVALID, "%s: short line: %s", path, l.c_str());
      const int pc = pnt_code(t[1]), lc = lig_code(t[2]);
      if (pc < 0 || lc < 0) return e.fail(GD_ERR_INVALID, "%s: unknown type in: %s", path, l.c_str());
      tab[pc][lc][0] = fnum(t[3]);
      tab[pc][lc][1] = fnum(t[4]);
      return GD_OK;
    };
    auto vec10 = [&](float *dst) -> gd_status {
      if (t.size() != GD_MAXWEI) return e.fail(GD_ERR_INVALID, "%s: %s needs 9 values", path, k.c_str());
      for (int i = 0; i + 1 < GD_MAXWEI; ++i) dst[i] = fnum(t[i + 1]);
      return GD_OK;
    };
    gd_status s = GD_OK;
    if (k == "VDW") s = pair_table(p.vdw);
    else if (k == "PMF") s = pair_table(p.pmf);
    else if (k == "HDB") s = pair_table(p.hdb);
    else if (k == "WEI") s = vec10(p.w);
    else if (k == "NOA") s = vec10(p.a);
    else if (k == "NOB") s = vec10(p.b);
    else if (k == "PLJ" && t.size() >= 4) {
      for (int i = 0; i < 3; ++i) p.lj[i] = fnum(t[i + 1]);
    } else if (k == "PEL" && t.size() >= 3) {
      for (int i = 0; i < 2; ++i) p.el[i] = fnum(t[i + 1]);
    } else if (k == "ELE" && t.size() >= 3) {
      // a point name, or a one-letter residue + 'C' for the CA point of that residue
      int idx;
      if (t[1].size() >= 2 && t[1][1] == 'C') idx = res1_code(t[1][0]) < 0 ? -1 : res1_code(t[1][0]) + 30;
      else idx = pnt_code(t[1]);
      if (idx < 0) return e.fail(GD_ERR_INVALID, "%s: unknown ELE point: %s", path, l.c_str());
      p.ele[idx] = fnum(t[2]);
    } else if (k == "HPP" && t.size() >= 3) {
      const int r = res1_code(t[1].empty() ? 0 : t[1][0]);
      if (r < 0) return e.fail(GD_ERR_INVALID, "%s: unknown residue: %s", path, l.c_str());
      p.hpp[r] = fnum(t[2]);
    } else if (k == "HPL" && t.size() >= 4) {
      const int lc = lig_code(t[1]);
      if (lc < 0) return e.fail(GD_ERR_INVALID, "%s: unknown ligand type: %s", path, l.c_str());
      p.hpl[lc][0] = fnum(t[2]);
      p.hpl[lc][1] = fnum(t[3]);
    } else if (k == "KDE" && t.size() >= 2) {
      p.kde = fnum(t[1]);
    }
    if (s != GD_OK) return s;
  }
  // hydrophobicity rescaled to [-1, 1] (src/load.C:1636-1659)
  float lo = 1e6f, hi = -1e6f;
  for (float v : p.hpp) {
    lo = std::min(lo, v);
    hi = std::max(hi, v);
  }
  for (float &v : p.hpp) {
    v = ((v - lo) / (hi - lo)) * 2.0f - 1.0f;
    v = std::min(1.0f, std::max(-1.0f, v));
  }
  return GD_OK;
}

// one value per line, first 30 characters (loadWeight / loadNorPara)
gd_status load_column(const char *path, float *dst, int n, Err &e) {
  std::vector<std::string> lines;
  if (!read_lines(path, lines)) return e.fail(GD_ERR_INVALID, "cannot open %s", path);
  for (int i = 0; i < (int)lines.size() && i < n; ++i) {
    std::istringstream ss(lines[i].substr(0, 30));
    float v = 0.0f;
    ss >> v;
    dst[i] = v;
  }
  return GD_OK;
}

// OptimizeEnepara, src/util.C:717-768
void derive_enepara(const RawPara &p, gd_enepara &ep) {
  std::memset(&ep, 0, sizeof ep);
  const float sqrt_2_pi = sqrtf(2.0f * 3.1415926535f);
  for (int lt = 0; lt < GD_MAXTP2; ++lt)
    for (int pt = 0; pt < GD_MAXTP1; ++pt) {
      const float s = p.vdw[pt][lt][0] * p.lj[2];
      ep.p1a[lt][pt] = 2.0f * p.vdw[pt][lt][1] * powf(s, 9.0f);
      ep.p2a[lt][pt] = 3.0f * p.vdw[pt][lt][1] * powf(s, 6.0f);
      ep.pmf0[lt][pt] = p.pmf[pt][lt][0];
      ep.pmf1[lt][pt] = p.pmf[pt][lt][1];
      ep.hdb0[lt][pt] = p.hdb[pt][lt][0];
      ep.hdb1[lt][pt] = 1.0f / p.hdb[pt][lt][1];
    }
  ep.lj0 = p.lj[0];
  ep.lj1 = p.lj[1];
  ep.el0 = p.el[0];
  ep.a1 = 4.0f - 3.0f * p.el[0];
  ep.b1 = 2.0f * p.el[0] - 3.0f;
  ep.el1 = p.el[1];
  for (int i = 0; i < GD_MAXTP3; ++i) ep.ele[i] = p.ele[i];
  for (int i = 0; i < GD_MAXTP4; ++i) ep.hpp[i] = p.hpp[i];
  for (int i = 0; i < GD_MAXTP2; ++i) {
    ep.hpl0[i] = p.hpl[i][0];
    ep.hpl1[i] = p.hpl[i][1];
    ep.hpl2[i] = logf(1.0f / (ep.hpl1[i] * sqrt_2_pi));
  }
  ep.kde2 = -0.5f / (p.kde * p.kde);
  ep.kde3 = powf(p.kde * sqrt_2_pi, 3.0f);
  for (int i = 0; i < GD_MAXWEI; ++i) {
    ep.w[i] = p.w[i];
    ep.a_para[i] = p.a[i];
    ep.b_para[i] = p.b[i];
  }
}

// ---- ligand SDF -------------------------------------------------------------
struct LigConf {
  std::vector<float> x, y, z;
};

struct LigandData {
  int sections = 0, ens_total = 0, lna = 0, lnb = 0;
  bool typed = false;
  std::vector<int> t, n;
  std::vector<float> c;
  LigConf def;                // the record's atom block
  std::vector<LigConf> ens;   // every ENSEMBLE_COORDS line, in file order
  std::vector<float> rmsd;    // ENSEMBLE_RMSD, one per line
  std::vector<LigConf> conf;  // docked conformers: def + ens with RMSD > 0.1
};

// One SDF record: readLigandSections, writeLigProperty, writeLigAtomProperty,
// writeDefaultLigAtomCoord, getLigEnsemble{Total,Coords,Rmsd}, writeEnsAtomCoord
// and loadOneLigand's RMSD filter (src/load.C:157-447).
gd_status load_sdf(const char *path, int record, bool need_types, LigandData &L, Err &e) {
  std::vector<std::string> lines;
  if (!read_lines(path, lines)) return e.fail(GD_ERR_INVALID, "cannot open %s", path);
  // records end with "$$$$"; the first one is docked
  std::vector<std::string> sect;
  for (const std::string &l : lines) {
    if (L.sections == record) sect.push_back(l);
    if (l == "$$$$") L.sections += 1;
  }
  if (record < 0 || L.sections <= record || sect.size() < 5)
    return e.fail(GD_ERR_INVALID, "%s: no $$$$-terminated record %d (%d records)", path, record,
                  L.sections);
  const int lna = icol(sect[3], 0, 3), lnb = icol(sect[3], 3, 3);
  if (lna <= 0 || lna > GD_MAX_LNA || lnb < 0 || (int)sect.size() < 4 + lna + lnb)
    return e.fail(GD_ERR_INVALID, "%s: bad atom count %d", path, lna);
  L.lna = lna;
  L.lnb = lnb;
  L.t.assign(lna, -1);
  L.c.assign(lna, 0.0f);
  L.n.resize(lna);
  for (int i = 0; i < lna; ++i) {
    const std::string &a = sect[4 + i];
    L.def.x.push_back(fcol(a, 0, 10));
    L.def.y.push_back(fcol(a, 10, 10));
    L.def.z.push_back(fcol(a, 20, 10));
    L.n[i] = i;  // atom number: the MCS fields are indexed by number + 1
  }
  std::vector<std::string> coords;
  for (size_t i = 4 + lna + lnb; i < sect.size(); ++i) {
    const std::string &l = sect[i];
    if (l.find("OB_ATOM_TYPES") != std::string::npos && i + 1 < sect.size()) {
      const auto tk = tokens(sect[i + 1]);
      for (int a = 0; a < lna && a < (int)tk.size(); ++a) L.t[a] = lig_code(tk[a]);
      L.typed = true;
    } else if ((l.find("OB_ATOMIC_CHARGES") != std::string::npos ||
                l.find("OB_CHARGES") != std::string::npos) && i + 1 < sect.size()) {
      const auto tk = tokens(sect[i + 1]);
      for (int a = 0; a < lna && a < (int)tk.size(); ++a) L.c[a] = fnum(tk[a]);
    } else if (l.find("ENSEMBLE_TOTAL") != std::string::npos && i + 1 < sect.size()) {
      L.ens_total = std::atoi(sect[i + 1].c_str());
    } else if (l.find("ENSEMBLE_RMSD") != std::string::npos) {
      for (size_t j = i + 1; j < sect.size() && !sect[j].empty() && std::isdigit((unsigned char)sect[j][0]); ++j)
        for (const auto &v : tokens(sect[j])) L.rmsd.push_back(fnum(v));
    } else if (l.find("ENSEMBLE_COORDS") != std::string::npos) {
      for (size_t j = i + 1; j < sect.size() && !sect[j].empty(); ++j) coords.push_back(sect[j]);
    }
  }
  if (need_types)
    for (int a = 0; a < lna; ++a)
      if (L.t[a] < 0)
        return e.fail(GD_ERR_INVALID, "%s: atom %d has no known OB_ATOM_TYPES entry", path, a + 1);
  if (L.rmsd.size() != coords.size())
    return e.fail(GD_ERR_INVALID, "%s: %zu ENSEMBLE_RMSD values for %zu ENSEMBLE_COORDS lines", path,
                  L.rmsd.size(), coords.size());
  for (size_t k = 0; k < coords.size(); ++k) {
    const auto v = tokens(coords[k]);
    if ((int)v.size() < 3 * lna)
      return e.fail(GD_ERR_INVALID, "%s: ensemble conformer %zu has %zu values", path, k, v.size());
    LigConf cf;
    for (int a = 0; a < lna; ++a) {
      cf.x.push_back(fnum(v[3 * a]));
      cf.y.push_back(fnum(v[3 * a + 1]));
      cf.z.push_back(fnum(v[3 * a + 2]));
    }
    L.ens.push_back(cf);
  }
  L.conf.push_back(L.def);
  for (size_t k = 0; k < L.ens.size(); ++k)
    if (L.rmsd[k] > 0.1f) L.conf.push_back(L.ens[k]);  // MINLIGRMSD, src/size.h:82
  return GD_OK;
}

// moveLigand2ItsCenterFrame, src/load.C:239-260 (FP32 sums in atom order)
void centre(LigConf &cf) {
  const int n = (int)cf.x.size();
  float sx = 0.f, sy = 0.f, sz = 0.f;
  for (int i = 0; i < n; ++i) {
    sx += cf.x[i];
    sy += cf.y[i];
    sz += cf.z[i];
  }
  const float cx = sx / n, cy = sy / n, cz = sz / n;
  for (int i = 0; i < n; ++i) {
    cf.x[i] -= cx;
    cf.y[i] -= cy;
    cf.z[i] -= cz;
  }
}

// ---- protein PDB -> coarse points -------------------------------------------
struct Point {
  int r, t, d, c;  // residue index, point type, residue code, class
};

struct ProteinData {
  int n_res = 0;
  std::vector<int> seq3;  // residue serial numbers
  std::vector<Point> pts;
  std::vector<std::vector<float>> x, y, z;  // [conformer][point]
};

gd_status load_pdb(const char *path, ProteinData &P, Err &e) {
  std::vector<std::string> lines;
  if (!read_lines(path, lines)) return e.fail(GD_ERR_INVALID, "cannot open %s", path);
  // points from the first model's CA records (src/load.C:962-1058)
  for (const std::string &l : lines) {
    if (l.size() > 53) {
      if (!starts(l, "ATOM  ") || l.substr(12, 4) != " CA ") continue;
      const int rc = res3_code(l.substr(17, 3));
      if (rc < 0) return e.fail(GD_ERR_INVALID, "%s: unknown residue %s", path, l.substr(17, 3).c_str());
      const int r = P.n_res;
      P.seq3.push_back(icol(l, 22, 4));
      P.pts.push_back({r, 0, rc, 0});
      if (r > 1) P.pts.push_back({r - 1, 1, rc, 1});
      for (int s = 0; s < kSide[rc].n; ++s) P.pts.push_back({r, kSide[rc].t[s], rc, kSide[rc].c[s]});
      P.n_res += 1;
    } else if (starts(l, "ENDMDL")) {
      break;
    }
  }
  if (P.pts.empty()) return e.fail(GD_ERR_INVALID, "%s: no CA atoms", path);
  // coordinates: per point, the mean of its atoms within each block that
  // ends at ENDMDL or TER (one block per protein conformer)
  struct Atom {
    int ser;
    std::string name;
    float x, y, z;
  };
  std::vector<std::vector<Atom>> blocks(1);
  for (const std::string &l : lines) {
    if (l.size() > 53) {
      if (starts(l, "ATOM  "))
        blocks.back().push_back({icol(l, 22, 4), l.substr(12, 4), fcol(l, 30, 8), fcol(l, 38, 8),
                                 fcol(l, 46, 8)});
    } else if (starts(l, "ENDMDL") || starts(l, "TER")) {
      blocks.emplace_back();
    }
  }
  blocks.pop_back();  // atoms after the last terminator belong to no conformer
  // a terminator right after another (TER then ENDMDL) closes an empty block:
  // it has no coordinates to give a conformer, so it is not one
  blocks.erase(std::remove_if(blocks.begin(), blocks.end(),
                              [](const std::vector<Atom> &b) { return b.empty(); }),
               blocks.end());
  if (blocks.empty()) return e.fail(GD_ERR_INVALID, "%s: no ENDMDL or TER record", path);
  const int np = (int)P.pts.size();
  P.x.assign(blocks.size(), std::vector<float>(np, 0.0f));
  P.y = P.x;
  P.z = P.x;
  for (size_t b = 0; b < blocks.size(); ++b)
    for (int i = 0; i < np; ++i) {
      const Point &pt = P.pts[i];
      const int s3 = P.seq3[pt.r];
      const int s6 = pt.r > 0 ? P.seq3[pt.r - 1] : -1;
      const PointRule rule = rule_of(pt.t);
      float sx = 0.f, sy = 0.f, sz = 0.f, sn = 0.f;
      for (const Atom &a : blocks[b]) {
        bool take = false;
        const std::string &nm = a.name;
        const bool bb = nm == " N  " || nm == " CA " || nm == " C  " || nm == " O  ";
        if (a.ser == s3) {
          if (rule == R_CA) take = nm == " CA ";
          else if (rule == R_SIDE) take = !bb;
          else if (rule == R_DISTAL) take = !bb && nm != " CB " && nm != " CG ";
          else if (rule == R_PROX) take = nm == " CB " || nm == " CG ";
        }
        if (rule == R_PEPTIDE && pt.r > 0 && pt.r < P.n_res - 1)
          take = take || (a.ser == s3 && nm == " N  ") ||
                 (a.ser == s6 && (nm == " C  " || nm == " O  "));
        if (take) {
          sx += a.x;
          sy += a.y;
          sz += a.z;
          sn += 1.0f;
        }
      }
      if (sn > 0.0f) {
        P.x[b][i] = sx / sn;
        P.y[b][i] = sy / sn;
        P.z[b][i] = sz / sn;
      }
    }
  return GD_OK;
}

// ---- force field (.ff) --------------------------------------------------------
struct LhmData {
  bool has_center = false;
  float center[3] = {0, 0, 0};
  std::vector<float> kx, ky, kz;
  std::vector<int> kt;
  int kde_dropped = 0;
  std::vector<int> psp_res, psp_t;
  std::vector<float> psp_v;
  struct Mcs {
    float tcc = 0.f;
    std::vector<int> idx;
    std::vector<float> x, y, z;
  };
  std::vector<Mcs> mcs;
};

bool is_number(const std::string &s) {
  if (s.empty()) return false;
  char *end = nullptr;
  std::strtof(s.c_str(), &end);
  return end && *end == 0;
}

// Sanity bounds on indices read from files (they size host tables).
constexpr int kMaxResidue = 99999;
constexpr int kMaxMcsField = 4095;

gd_status load_ff(const char *path, const char *complex_id, LhmData &H, Err &e) {
  std::vector<std::string> lines;
  if (!read_lines(path, lines)) return e.fail(GD_ERR_INVALID, "cannot open %s", path);
  const std::string id = complex_id ? complex_id : "";
  for (const std::string &l : lines) {
    if (l.find("CENTER") != std::string::npos) {
      const auto t = tokens(l);
      if (t.size() >= 4) {
        for (int i = 0; i < 3; ++i) H.center[i] = fnum(t[i + 1]);
        H.has_center = true;
      }
    }
    if (l.size() <= 3) continue;
    const auto t = tokens(l);
    if (t.empty()) continue;
    if (t[0] == "KDE" && t.size() >= 5) {
      const int c = lig_code(t[1]);
      if (c < 0) {
        H.kde_dropped += 1;  // the reference keeps it with a type no atom has
        continue;
      }
      H.kt.push_back(c);
      H.kx.push_back(fnum(t[2]));
      H.ky.push_back(fnum(t[3]));
      H.kz.push_back(fnum(t[4]));
    } else if (t[0] == "PSP" && t.size() >= 4) {
      const int c = lig_code(t[2]);
      if (c < 0) continue;
      const int r = std::atoi(t[1].c_str());
      if (r < 0 || r > kMaxResidue)
        return e.fail(GD_ERR_INVALID, "%s: PSP residue %d outside [0, %d]", path, r, kMaxResidue);
      H.psp_res.push_back(r);
      H.psp_t.push_back(c);
      H.psp_v.push_back(fnum(t[3]));
    } else if (t[0] == "MCS" && t.size() >= 4) {
      // "MCS <id> tcc n (i x y z)*" when the record names this complex,
      // else one more leading field (src/load.C:1420-1460)
      const bool own = t[1] == id && is_number(t[2]);
      const size_t o = own ? 2 : 3;
      if (t.size() < o + 2) return e.fail(GD_ERR_INVALID, "%s: short MCS record", path);
      LhmData::Mcs m;
      m.tcc = fnum(t[o]);
      const int n = std::atoi(t[o + 1].c_str());
      if ((int)t.size() < (int)(o + 2 + 4 * n))
        return e.fail(GD_ERR_INVALID, "%s: MCS record with %d fields is short", path, n);
      for (int k = 0; k < n; ++k) {
        const int ai = std::atoi(t[o + 2 + 4 * k].c_str());
        if (ai < -1 || ai > kMaxMcsField)
          return e.fail(GD_ERR_INVALID, "%s: MCS field %d outside [-1, %d]", path, ai, kMaxMcsField);
        m.idx.push_back(ai);
        m.x.push_back(fnum(t[o + 3 + 4 * k]));
        m.y.push_back(fnum(t[o + 4 + 4 * k]));
        m.z.push_back(fnum(t[o + 5 + 4 * k]));
      }
      H.mcs.push_back(m);
    }
  }
  if (!H.has_center) return e.fail(GD_ERR_INVALID, "%s: no CENTER record", path);
  return GD_OK;
}

}  // namespace

struct gd_loaded {
  gd_complex_desc d{};
  gd_enepara ep{};
  int sections = 0, ens_total = 0, residues = 0, kde_dropped = 0;
  std::vector<float> lig_x, lig_y, lig_z, lig_c;
  std::vector<int> lig_t, lig_n;
  std::vector<float> prt_x, prt_y, prt_z, prt_ele, prt_hpp;
  std::vector<int> prt_t, prt_c, prt_seq3r, prt_flag;
  std::vector<float> psp;
  std::vector<float> kde_x, kde_y, kde_z;
  std::vector<int> kde_t;
  std::vector<float> mcs_x, mcs_y, mcs_z, mcs_tcc;
  std::vector<gd_temp> temp;
};

extern "C" {

namespace {
gd_status load_complex_impl(const gd_input_files *files, const gd_dock_params *params,
                            gd_loaded **out, Err &e);
gd_status sdf_read_impl(const char *path, int record, int centre_each, gd_sdf_record *info,
                        float *xyz, int max_conf, float *rmsd, int max_rmsd, int *types,
                        float *charges, Err &e);
gd_status write_pose_sdf_impl(const char *ligand_sdf, int lna, const float *x, const float *y,
                              const float *z, const char *out_path, Err &e);
}  // namespace

// The loaders allocate host tables sized by the files; an allocation failure
// or a library exception becomes a status, never an exception across the ABI.
#define GD_IO_GUARD(call)                                                   \
  try {                                                                     \
    return call;                                                            \
  } catch (const std::bad_alloc &) {                                        \
    return e.fail(GD_ERR_NOMEM, "out of host memory");                      \
  } catch (const std::exception &x) {                                       \
    return e.fail(GD_ERR_INVALID, "input error: %s", x.what());             \
  }

gd_status gd_load_complex(const gd_input_files *files, const gd_dock_params *params,
                          gd_loaded **out, char *err, size_t err_len) {
  Err e{err, err_len};
  if (err && err_len) err[0] = 0;
  if (out) *out = nullptr;
  GD_IO_GUARD(load_complex_impl(files, params, out, e))
}

}  // extern "C"

namespace {
gd_status load_complex_impl(const gd_input_files *files, const gd_dock_params *params,
                            gd_loaded **out, Err &e) {
  if (!files || !params || !out || !files->ligand_sdf || !files->protein_pdb || !files->lhm_ff ||
      !files->enepara)
    return e.fail(GD_ERR_INVALID, "missing input file path");
  if (params->n_tmp < 1 || params->n_tmp > GD_MAX_NTMP || !(params->floor_temp > 0.0f) ||
      (params->n_tmp > 1 && !(params->ceiling_temp > 0.0f)))
    return e.fail(GD_ERR_INVALID, "bad temperature parameters");
  *out = nullptr;
  RawPara raw;
  LigandData L;
  ProteinData P;
  LhmData H;
  gd_status s;
  if ((s = load_paras(files->enepara, raw, e)) != GD_OK) return s;
  if (files->weight && (s = load_column(files->weight, raw.w, GD_MAXWEI, e)) != GD_OK) return s;
  if (files->nor_a && (s = load_column(files->nor_a, raw.a, GD_MAXWEI, e)) != GD_OK) return s;
  if (files->nor_b && (s = load_column(files->nor_b, raw.b, GD_MAXWEI, e)) != GD_OK) return s;
  if ((s = load_sdf(files->ligand_sdf, 0, true, L, e)) != GD_OK) return s;
  if ((s = load_pdb(files->protein_pdb, P, e)) != GD_OK) return s;
  if ((s = load_ff(files->lhm_ff, files->complex_id, H, e)) != GD_OK) return s;

  std::unique_ptr<gd_loaded> owner(new gd_loaded);
  gd_loaded *g = owner.get();
  g->sections = L.sections;
  g->ens_total = L.ens_total;
  g->residues = P.n_res;
  g->kde_dropped = H.kde_dropped;
  derive_enepara(raw, g->ep);

  // ligand conformers in their own centroid frames
  const int lna = L.lna, n_lig = (int)L.conf.size();
  for (LigConf &cf : L.conf) centre(cf);
  for (const LigConf &cf : L.conf) {
    g->lig_x.insert(g->lig_x.end(), cf.x.begin(), cf.x.end());
    g->lig_y.insert(g->lig_y.end(), cf.y.begin(), cf.y.end());
    g->lig_z.insert(g->lig_z.end(), cf.z.begin(), cf.z.end());
  }
  g->lig_t = L.t;
  g->lig_c = L.c;
  g->lig_n = L.n;

  // receptor points (CopyProteinResidue, src/util.C:574-597)
  const int pnp = (int)P.pts.size(), n_prt = (int)P.x.size();
  int max_seq = 0;
  for (const Point &pt : P.pts) {
    g->prt_t.push_back(pt.t);
    g->prt_c.push_back(pt.c);
    const int dt = pt.t == 0 ? pt.d + 30 : pt.t;
    g->prt_ele.push_back(g->ep.ele[dt]);
    const int sr = P.seq3[pt.r];
    g->prt_seq3r.push_back(sr);
    max_seq = std::max(max_seq, sr);
    g->prt_flag.push_back((pt.c == 0 && pt.d == 12) || pt.c == 2 ? 1 : 0);
    g->prt_hpp.push_back(g->ep.hpp[pt.d]);
  }
  for (int c = 0; c < n_prt; ++c) {
    g->prt_x.insert(g->prt_x.end(), P.x[c].begin(), P.x[c].end());
    g->prt_y.insert(g->prt_y.end(), P.y[c].begin(), P.y[c].end());
    g->prt_z.insert(g->prt_z.end(), P.z[c].begin(), P.z[c].end());
  }
  for (int s3 : g->prt_seq3r)
    if (s3 < 0 || s3 > kMaxResidue)
      return e.fail(GD_ERR_INVALID, "residue serial number %d outside [0, %d]", s3, kMaxResidue);

  // PSP table [ligand type][residue serial]
  int n_res = max_seq + 1;
  for (int r : H.psp_res) n_res = std::max(n_res, r + 1);
  g->psp.assign((size_t)GD_MAXTP2 * n_res, 0.0f);
  for (size_t k = 0; k < H.psp_res.size(); ++k)
    if (H.psp_res[k] >= 0) g->psp[(size_t)H.psp_t[k] * n_res + H.psp_res[k]] = H.psp_v[k];

  g->kde_x = H.kx;
  g->kde_y = H.ky;
  g->kde_z = H.kz;
  g->kde_t = H.kt;

  // MCS templates: fields indexed by atom number + 1
  int n_mcs = lna + 1;
  for (const auto &m : H.mcs)
    for (int i : m.idx) n_mcs = std::max(n_mcs, i + 1);
  const int pos = (int)H.mcs.size();
  g->mcs_x.assign((size_t)pos * n_mcs, GD_MCS_INVALID_COORD);
  g->mcs_y = g->mcs_x;
  g->mcs_z = g->mcs_x;
  for (int m = 0; m < pos; ++m) {
    g->mcs_tcc.push_back(H.mcs[m].tcc);
    for (size_t k = 0; k < H.mcs[m].idx.size(); ++k) {
      const int i = H.mcs[m].idx[k];
      if (i < 0) continue;
      g->mcs_x[(size_t)m * n_mcs + i] = H.mcs[m].x[k];
      g->mcs_y[(size_t)m * n_mcs + i] = H.mcs[m].y[k];
      g->mcs_z[(size_t)m * n_mcs + i] = H.mcs[m].z[k];
    }
  }

  // SetTemperature, src/util.C:797-837 (k_B = 1)
  const int nt = params->n_tmp;
  g->temp.resize(nt);
  if (nt == 1) {
    g->temp[0] = {params->floor_temp, 0.0f - 1.0f / params->floor_temp, 0};
  } else {
    const float ratio = (float)std::exp(std::log(params->ceiling_temp / params->floor_temp) /
                                        (float)(nt - 1));
    float a = params->floor_temp;
    for (int i = 0; i < nt; ++i) {
      g->temp[i] = {a, 0.0f - 1.0f / a, i};
      a *= ratio;
    }
  }

  gd_complex_desc &d = g->d;
  d.n_lig = n_lig;
  d.n_prt = n_prt;
  d.n_tmp = nt;
  d.lna = lna;
  d.pnp = pnp;
  d.pnk = (int)g->kde_t.size();
  d.pos = pos;
  d.n_res = n_res;
  d.n_mcs = n_mcs;
  d.lig_x = g->lig_x.data();
  d.lig_y = g->lig_y.data();
  d.lig_z = g->lig_z.data();
  d.lig_t = g->lig_t.data();
  d.lig_c = g->lig_c.data();
  d.lig_n = g->lig_n.data();
  for (int i = 0; i < 3; ++i) d.pocket_center[i] = H.center[i];
  d.prt_x = g->prt_x.data();
  d.prt_y = g->prt_y.data();
  d.prt_z = g->prt_z.data();
  d.prt_t = g->prt_t.data();
  d.prt_c = g->prt_c.data();
  d.prt_ele = g->prt_ele.data();
  d.prt_seq3r = g->prt_seq3r.data();
  d.prt_c0_and_d12_or_c2 = g->prt_flag.data();
  d.prt_hpp = g->prt_hpp.data();
  d.psp = g->psp.data();
  d.kde_x = g->kde_x.data();
  d.kde_y = g->kde_y.data();
  d.kde_z = g->kde_z.data();
  d.kde_t = g->kde_t.data();
  d.mcs_x = g->mcs_x.data();
  d.mcs_y = g->mcs_y.data();
  d.mcs_z = g->mcs_z.data();
  d.mcs_tcc = g->mcs_tcc.data();
  d.enepara = &g->ep;
  d.temp = g->temp.data();
  for (int i = 0; i < 3; ++i) d.move_scale[i] = params->ts;
  for (int i = 3; i < 6; ++i) d.move_scale[i] = params->rs;
  d.move_kind = params->move_kind;
  *out = owner.release();
  return GD_OK;
}
}  // namespace

extern "C" {

const gd_complex_desc *gd_loaded_desc(const gd_loaded *l) { return l ? &l->d : nullptr; }

void gd_loaded_info(const gd_loaded *l, int *sdf_sections, int *ensemble_total, int *residues,
                    int *kde_dropped) {
  if (!l) return;
  if (sdf_sections) *sdf_sections = l->sections;
  if (ensemble_total) *ensemble_total = l->ens_total;
  if (residues) *residues = l->residues;
  if (kde_dropped) *kde_dropped = l->kde_dropped;
}

void gd_loaded_free(gd_loaded *l) { delete l; }

gd_status gd_sdf_read(const char *path, int record, int centre_each, gd_sdf_record *info,
                      float *xyz, int max_conf, float *rmsd, int max_rmsd, int *types,
                      float *charges, char *err, size_t err_len) {
  Err e{err, err_len};
  if (err && err_len) err[0] = 0;
  GD_IO_GUARD(sdf_read_impl(path, record, centre_each, info, xyz, max_conf, rmsd, max_rmsd, types,
                            charges, e))
}

}  // extern "C"

namespace {
gd_status sdf_read_impl(const char *path, int record, int centre_each, gd_sdf_record *info,
                        float *xyz, int max_conf, float *rmsd, int max_rmsd, int *types,
                        float *charges, Err &e) {
  if (!path || !info) return e.fail(GD_ERR_INVALID, "missing path or info");
  LigandData L;
  gd_status s;
  if ((s = load_sdf(path, record, false, L, e)) != GD_OK) return s;
  info->records = L.sections;
  info->lna = L.lna;
  info->lnb = L.lnb;
  info->typed = L.typed ? 1 : 0;
  info->ensemble_total = L.ens_total;
  info->ensemble_lines = (int)L.ens.size();
  info->conformers = (int)L.conf.size();
  const int lna = L.lna;
  for (int k = 0; xyz && k < max_conf && k <= (int)L.ens.size(); ++k) {
    LigConf cf = k == 0 ? L.def : L.ens[k - 1];
    if (centre_each) centre(cf);
    float *o = xyz + (size_t)k * 3 * lna;
    std::copy(cf.x.begin(), cf.x.end(), o);
    std::copy(cf.y.begin(), cf.y.end(), o + lna);
    std::copy(cf.z.begin(), cf.z.end(), o + 2 * lna);
  }
  for (int k = 0; rmsd && k < max_rmsd && k < (int)L.rmsd.size(); ++k) rmsd[k] = L.rmsd[k];
  if (types) std::copy(L.t.begin(), L.t.end(), types);
  if (charges) std::copy(L.c.begin(), L.c.end(), charges);
  return GD_OK;
}
}  // namespace

extern "C" {

// ---- trajectory output --------------------------------------------------------

gd_status gd_write_states_csv(const char *path, const gd_record *recs, int64_t n, char *err,
                              size_t err_len) {
  Err e{err, err_len};
  if (err && err_len) err[0] = 0;
  if (!path || (n > 0 && !recs) || n < 0) return e.fail(GD_ERR_INVALID, "bad arguments");
  FILE *f = std::fopen(path, "w");
  if (!f) return e.fail(GD_ERR_INVALID, "cannot write %s", path);
  std::fputs("lig prt cms rmsd mv1 mv2 mv3 mv4 mv5 mv6 "
             "vdw ele pmf psp hdb hpc kde lhm edst ener\n", f);
  for (int64_t i = 0; i < n; ++i) {
    const gd_record &r = recs[i];
    std::fprintf(f, "%d %d %.4f %.4f", r.replica.idx_lig, r.replica.idx_prt, r.energy.cms,
                 r.energy.rmsd);
    for (int j = 0; j < 6; ++j) std::fprintf(f, " %.4f", r.movematrix[j]);
    for (int j = 0; j < GD_MAXWEI; ++j) std::fprintf(f, " %.4f", r.energy.e[j]);
    std::fputc('\n', f);
  }
  const bool ok = std::fclose(f) == 0;
  return ok ? GD_OK : e.fail(GD_ERR_INVALID, "error writing %s", path);
}

gd_status gd_write_medoids_csv(const char *path, const gd_record *recs, const int *cluster_size,
                               int64_t n, char *err, size_t err_len) {
  Err e{err, err_len};
  if (err && err_len) err[0] = 0;
  if (!path || (n > 0 && (!recs || !cluster_size)) || n < 0)
    return e.fail(GD_ERR_INVALID, "bad arguments");
  FILE *f = std::fopen(path, "w");
  if (!f) return e.fail(GD_ERR_INVALID, "cannot write %s", path);
  std::fputs("lig prt t0 t1 t2 r0 r1 r2 vdw ele pmf psp hdb hpc kde lhm edst "
             "cms rmsd ener cluster_sz\n", f);
  for (int64_t i = 0; i < n; ++i) {
    const gd_record &r = recs[i];
    std::fprintf(f, "%d %d ", r.replica.idx_lig, r.replica.idx_prt);
    for (int j = 0; j < 6; ++j) std::fprintf(f, "%f ", r.movematrix[j]);
    for (int j = 0; j < GD_MAXWEI - 1; ++j) std::fprintf(f, "%f ", r.energy.e[j]);
    std::fprintf(f, "%f %f %f %d \n", r.energy.cms, r.energy.rmsd, r.energy.e[GD_MAXWEI - 1],
                 cluster_size[i]);
  }
  const bool ok = std::fclose(f) == 0;
  return ok ? GD_OK : e.fail(GD_ERR_INVALID, "error writing %s", path);
}

// ---- flat record file -----------------------------------------------------------
// 64-byte header, then gd_record entries (little-endian, 92 bytes each) in
// append order: the content of the reference's per-dump LigRecord datasets
// (src/hdf5_datastruct_ligrecord_create.h:53-84) without the padding of the
// fixed STEPS_PER_DUMP slots.
namespace {
struct RecHeader {
  char magic[8];       // "GDREC001"
  uint32_t rec_size;   // sizeof(gd_record)
  uint32_t version;    // 1
  uint64_t n_records;
  char pad[40];
};
static_assert(sizeof(RecHeader) == 64, "record file header is 64 bytes");
static_assert(sizeof(gd_record) == 92, "gd_record is 92 bytes");

gd_status read_header(FILE *f, const char *path, RecHeader &h, Err &e) {
  if (std::fread(&h, sizeof h, 1, f) != 1) return e.fail(GD_ERR_INVALID, "%s: no header", path);
  if (std::memcmp(h.magic, "GDREC001", 8) != 0 || h.rec_size != sizeof(gd_record) || h.version != 1)
    return e.fail(GD_ERR_INVALID, "%s: not a record file of this layout", path);
  return GD_OK;
}
}  // namespace

gd_status gd_records_write(const char *path, int append, const gd_record *recs, int64_t n,
                           char *err, size_t err_len) {
  Err e{err, err_len};
  if (err && err_len) err[0] = 0;
  if (!path || n < 0 || (n > 0 && !recs)) return e.fail(GD_ERR_INVALID, "bad arguments");
  RecHeader h{};
  FILE *f = append ? std::fopen(path, "r+b") : nullptr;
  if (f) {
    gd_status s = read_header(f, path, h, e);
    if (s != GD_OK) {
      std::fclose(f);
      return s;
    }
  } else {
    f = std::fopen(path, "w+b");
    if (!f) return e.fail(GD_ERR_INVALID, "cannot write %s", path);
    std::memcpy(h.magic, "GDREC001", 8);
    h.rec_size = sizeof(gd_record);
    h.version = 1;
    h.n_records = 0;
  }
  bool ok = std::fseek(f, (long)(sizeof h + h.n_records * sizeof(gd_record)), SEEK_SET) == 0;
  ok = ok && (n == 0 || std::fwrite(recs, sizeof(gd_record), (size_t)n, f) == (size_t)n);
  h.n_records += (uint64_t)n;
  ok = ok && std::fseek(f, 0, SEEK_SET) == 0 && std::fwrite(&h, sizeof h, 1, f) == 1;
  ok = (std::fclose(f) == 0) && ok;
  return ok ? GD_OK : e.fail(GD_ERR_INVALID, "error writing %s", path);
}

gd_status gd_records_count(const char *path, int64_t *n, char *err, size_t err_len) {
  Err e{err, err_len};
  if (err && err_len) err[0] = 0;
  if (!path || !n) return e.fail(GD_ERR_INVALID, "bad arguments");
  FILE *f = std::fopen(path, "rb");
  if (!f) return e.fail(GD_ERR_INVALID, "cannot open %s", path);
  RecHeader h{};
  gd_status s = read_header(f, path, h, e);
  std::fclose(f);
  if (s == GD_OK) *n = (int64_t)h.n_records;
  return s;
}

gd_status gd_records_read(const char *path, int64_t first, int64_t n, gd_record *out, char *err,
                          size_t err_len) {
  Err e{err, err_len};
  if (err && err_len) err[0] = 0;
  if (!path || first < 0 || n < 0 || (n > 0 && !out)) return e.fail(GD_ERR_INVALID, "bad arguments");
  FILE *f = std::fopen(path, "rb");
  if (!f) return e.fail(GD_ERR_INVALID, "cannot open %s", path);
  RecHeader h{};
  gd_status s = read_header(f, path, h, e);
  if (s == GD_OK && (uint64_t)(first + n) > h.n_records)
    s = e.fail(GD_ERR_INVALID, "%s: records [%lld, %lld) past the %llu stored", path,
               (long long)first, (long long)(first + n), (unsigned long long)h.n_records);
  if (s == GD_OK) {
    const bool ok = std::fseek(f, (long)(sizeof h + (uint64_t)first * sizeof(gd_record)), SEEK_SET) == 0 &&
                    (n == 0 || std::fread(out, sizeof(gd_record), (size_t)n, f) == (size_t)n);
    if (!ok) s = e.fail(GD_ERR_INVALID, "%s: short read", path);
  }
  std::fclose(f);
  return s;
}

// ---- trace tool: a pose as ligand coordinates and as an SDF -----------------------

// PlaceLigand (src/util.C:456-497): the conformer (in its centroid frame)
// rotated by the Euler angles mm[3..5], then moved to the pocket centre plus
// mm[0..2]. FP32, as the reference's host code.
gd_status gd_pose_coords(const gd_complex_desc *d, int idx_lig, const float mm[6], float *x,
                         float *y, float *z) {
  if (!d || !mm || !x || !y || !z || idx_lig < 0 || idx_lig >= d->n_lig) return GD_ERR_INVALID;
  const float s1 = sinf(mm[3]), c1 = cosf(mm[3]);
  const float s2 = sinf(mm[4]), c2 = cosf(mm[4]);
  const float s3 = sinf(mm[5]), c3 = cosf(mm[5]);
  const float r00 = c1 * c2, r01 = c1 * s2 * s3 - c3 * s1, r02 = s1 * s3 + c1 * c3 * s2;
  const float r10 = c2 * s1, r11 = c1 * c3 + s1 * s2 * s3, r12 = c3 * s1 * s2 - c1 * s3;
  const float r20 = -s2, r21 = c2 * s3, r22 = c2 * c3;
  const size_t o = (size_t)idx_lig * d->lna;
  for (int l = 0; l < d->lna; ++l) {
    const float a = d->lig_x[o + l], b = d->lig_y[o + l], c = d->lig_z[o + l];
    x[l] = r00 * a + r01 * b + r02 * c + mm[0] + d->pocket_center[0];
    y[l] = r10 * a + r11 * b + r12 * c + mm[1] + d->pocket_center[1];
    z[l] = r20 * a + r21 * b + r22 * c + mm[2] + d->pocket_center[2];
  }
  return GD_OK;
}

// replaceLigandCoords (src/util.C:504-549): the SDF with lines 5..4+lna
// (the first record's atom block) starting with the new coordinates as
// %10.4f%10.4f%10.4f; every other line unchanged.
gd_status gd_write_pose_sdf(const char *ligand_sdf, int lna, const float *x, const float *y,
                            const float *z, const char *out_path, char *err, size_t err_len) {
  Err e{err, err_len};
  if (err && err_len) err[0] = 0;
  GD_IO_GUARD(write_pose_sdf_impl(ligand_sdf, lna, x, y, z, out_path, e))
}

}  // extern "C"

namespace {
gd_status write_pose_sdf_impl(const char *ligand_sdf, int lna, const float *x, const float *y,
                              const float *z, const char *out_path, Err &e) {
  if (!ligand_sdf || !out_path || !x || !y || !z || lna <= 0)
    return e.fail(GD_ERR_INVALID, "bad arguments");
  std::ifstream in(ligand_sdf);
  if (!in.is_open()) return e.fail(GD_ERR_INVALID, "cannot open %s", ligand_sdf);
  std::vector<std::string> lines;
  for (std::string l; std::getline(in, l);) lines.push_back(l);
  if ((int)lines.size() < 4 + lna) return e.fail(GD_ERR_INVALID, "%s: fewer than %d atom lines", ligand_sdf, lna);
  for (int a = 0; a < lna; ++a)
    if (lines[4 + a].size() < 30) return e.fail(GD_ERR_INVALID, "%s: atom line %d is short", ligand_sdf, a + 1);
  FILE *f = std::fopen(out_path, "w");
  if (!f) return e.fail(GD_ERR_INVALID, "cannot write %s", out_path);
  for (size_t i = 0; i < lines.size(); ++i) {
    if (i >= 4 && i < (size_t)(4 + lna)) {
      const int a = (int)i - 4;
      std::fprintf(f, "%10.4f%10.4f%10.4f%s\n", x[a], y[a], z[a], lines[i].c_str() + 30);
    } else {
      std::fprintf(f, "%s\n", lines[i].c_str());
    }
  }
  const bool ok = std::fclose(f) == 0;
  return ok ? GD_OK : e.fail(GD_ERR_INVALID, "error writing %s", out_path);
}
}  // namespace
