// gd_kernels.h -- data layout shared by the host side and the kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/geauxdock.h"

namespace gd {

constexpr int NLT = GD_MAXTP2;            // ligand atom types
constexpr int NPT = GD_MAXTP1;            // receptor point types
constexpr int NPAR = 12;                  // per (receptor type, ligand type) record
constexpr int PAR_FLOATS = NPT * NLT * NPAR;
constexpr int KMAX = GD_MAX_REFCONTACT;

// per (receptor type, ligand type) parameter record
// the record: the pair parameters, then values the pair loop derives from them
// (the contact radius squared, the LJ denominator's a0 and a1, the PMF
// exponent's B and C), precomputed with the pair loop's FP32 expressions
enum { P_P1A = 0, P_P2A, P_PMF0, P_PMF1, P_HA, P_HB, P_HC, P_NR2, P_A0, P_A1, P_PB, P_PC };

// Receptor point kinds after sorting: which optional terms a point can
// contribute. 0: neither, 1: hydrophobic contact (c0_and_d12_or_c2), 2: also
// pocket-specific potential (class 2).
enum { KIND_PLAIN = 0, KIND_HPC = 1, KIND_PSP = 2 };

// One complex on the device. Receptor points are stored sorted by
// (type, kind) so each chunk has one parameter record per ligand atom.
struct DevComplex {
  int lna, pnp, pnk, pos, n_res, n_lig, n_prt, n_tmp;
  int n_chunks, lpr, apl, move_kind;
  int ref_total;  // contacts of the reference contact map
  int cid;        // complex id (random stream id)
  int da_share;   // 1: split groups share the surrogate's KDE and MCS sums
  float pocket[3];
  float move_scale[6];
  float el0, el1, a1, b1, kde2l, kde3;
  float w[GD_MAXWEI], a_para[GD_MAXWEI], b_para[GD_MAXWEI];
  float hpl0[NLT], hpl1[NLT], hpl2[NLT];
  float minus_beta[GD_MAX_NTMP];
  const float4 *prt_xyze;  // [n_prt * pnp] x, y, z, ele (sorted order)
  const float *prt_hpp;    // [pnp] hpp where c0_and_d12_or_c2, else 0
  const int *prt_psp;      // [pnp] seq3r where class 2, else -1
  const int4 *chunks;      // [n_chunks] start, length, type, kind
  const float *params;     // [NPT * NLT * NPAR]
  const float *lig_orig;   // [n_lig][3][lna]
  const int *lig_t;        // [lna]
  const float *lig_c;      // [lna]
  const float4 *mcs;       // [pos * lna] x, y, z, valid
  const float *mcs_tcc;    // [pos]
  const float *mcs_cnt;    // [pos] valid atoms per template
  // per (ligand conformer, template), over the template's valid atoms:
  // {A, obar}, {mbar, M00}, {M01, M02, M10, M11}, {M12, M20, M21, M22} with
  // A = sum |o - obar|^2 + sum |m - mbar|^2 and M = sum (m - mbar)(o - obar)^T,
  // so that sum |R o + t - m|^2 = A + n |R obar + t - mbar|^2 - 2 <R, M>
  const float4 *mcsq;      // [n_lig * pos * 4]
  const float4 *kde;       // [pnk] sorted by type
  const int2 *kde_range;   // [NLT] begin, end
  const float *psp;        // [NLT * n_res]
  int2 *ref;               // [lna * KMAX] sorted point index, pmf0 bits
  int *ref_cnt;            // [lna]
  const int *prt_cls;      // [pnp] receptor class t mod 4 (sorted order), surrogate features
  // surrogate feature grid [n_prt][sn^3][32] at origin (sox, soy, soz), spacing sh
  const float *sgrid;
  float sox, soy, soz, sh;
  int sn;
  // network output tabulated on the same grid per ligand atom type:
  // stab[n_prt][NLT][sn^3] (null until tabulated)
  const float *stab;
};

// One workgroup's work: replicas (or poses) of one complex that share one
// protein conformer; slots are ordered ladder by ladder, temperature slot
// fastest, so slots [g*n_tmp, (g+1)*n_tmp) form exchange ladder g. The
// pack's receptor split: `ns` lane groups of `half` lanes (a multiple of 64)
// each evaluate all of the pack's replicas over their share of every receptor
// chunk.
struct Pack {
  int cx, prt, n_slots, n_groups, slot_off;
  int ns, half;
};

// MODE_SUREVAL: surrogate total of given poses (delayed-acceptance kernels)
enum { MODE_MC = 0, MODE_INIT = 1, MODE_EVAL = 2, MODE_SUREVAL = 3 };

struct KParams {
  const DevComplex *cxs;
  const Pack *packs;  // this launch's workgroups use packs[pack_base + blockIdx.x]
  int pack_base;
  const int *slots;
  // replica state (MODE_MC / MODE_INIT), indexed by global replica id
  float *mm;
  gd_energy *e_old;
  float *best_e;   // lowest total energy each replica has held, and its pose
  float *best_mm;
  int *idx_tmp;
  int *acs_mc;
  int *acs_x;
  const int *rep_flat;
  gd_record *rec;
  int *rec_cnt;
  int rec_cap;
  // poses (MODE_EVAL), indexed by slot value
  const float *ev_mm;
  const int *ev_lig;
  gd_energy *ev_out;
  int mode;
  int step0, n_steps, nc;
  const int *step_dev;  // when set, the first step is read here (graph replays)
  int *queue;           // when set, persistent workgroups take packs from this counter
  int n_queue;          // packs in this launch (persistent mode)
  int off_q;            // LDS word for the queue broadcast
  uint32_t k0, k1;
  // LDS layout (bytes)
  int off_hpp, off_psp, off_par, off_kde, off_kr, off_ch, off_xe, off_xi, off_xa, off_xo;
  // receptor split (per pack): the groups exchange partial sums through LDS
  // at off_xp (two parity buffers of xp_floats); `unsplit` runs every pack as
  // one group of the whole workgroup
  int off_xp, xp_floats, unsplit;
  int off_da;  // per-wave surrogate scratch (direct network evaluation)
  int tile;    // receptor points per LDS tile of tiled classes (0: receptor resident in LDS)
  int sur_tab;  // 1: surrogate from the tabulated network outputs
  float *ev_sur;  // MODE_SUREVAL output, indexed by slot value
  // delayed acceptance: k surrogate sub-steps per exact evaluation (0 = plain)
  int da_k;
  float sur_scale;
  const unsigned short *sW1T, *sW2T;
  const float *sb2, *sw3, *sb3;
};

// Host handle of the MC kernel instance for `APL` ligand atoms per lane,
// mode `da` (0 plain, 1 delayed acceptance with the network per atom, 2 with
// its tabulated outputs) and receptor tiling; defined in gd_mc_inst.hip, one
// object per APL.
template <int APL>
const void *mc_kernel_handle(int da, bool tiled);
template <> const void *mc_kernel_handle<1>(int da, bool tiled);
template <> const void *mc_kernel_handle<2>(int da, bool tiled);
template <> const void *mc_kernel_handle<3>(int da, bool tiled);
template <> const void *mc_kernel_handle<4>(int da, bool tiled);

}  // namespace gd
