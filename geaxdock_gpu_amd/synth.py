"""Seeded synthetic docking complexes for the benchmark workloads.

A complex holds the arrays the reference's Run() consumes after its loaders
and Optimize* passes (src/dock.C:104-181): ligand conformers in the ligand
frame, protein conformers with per-point attributes, the PSP table, KDE points,
MCS templates, energy parameters, the temperature ladder and move scales.

Choices the workload descriptions leave open (all fixed by the seed):
  * receptor point class c ~ U{0..3}, residue code d ~ U{0..19}, three points
    per residue (seq3r = p // 3); c0_and_d12_or_c2 and hpp follow
    CopyProteinResidue (src/util.C:574-597); the receptor partial charge
    N(0, 0.3) clipped to +-1 is used directly as the per-point ele value;
  * receptor point types are drawn from the 24 ligand-type range [0, 24);
  * protein conformer 0 is the unperturbed receptor, conformer k > 0 adds
    N(0, 0.5 A) per coordinate (that workload's rule);
  * the ligand chain keeps the most compact of 32 dihedral draws; further
    conformers resample every dihedral; every conformer is centred on its
    own centroid (ligand frame);
  * KDE point types are drawn from the ligand's own atom types;
  * MCS template m picks `mcs_atoms` ligand atoms at random, places them at
    the start pose plus N(0, 1 A), marks the others 728492 (invalid), and has
    tcc ~ U(0, 1];
  * 10 % of PSP entries are N(0, 1), the rest 0;
  * energy-parameter tables are drawn with the ranges of the reference's
    data/parameters/paras (VDW sigma, epsilon; PMF; HDB for 13 % of type
    pairs; HPP; HPL) and derived as OptimizeEnepara does
    (src/util.C:717-768, HPP rescale src/load.C:1637-1658). The scalar lines
    PEL, PLJ, KDE and the WEI/NOA/NOB weights are the values of
    data/parameters/paras (GeauxDock, LA-SiGMA Application License 1.0).
"""
import ctypes as C
import math
import os

import numpy as np

from . import abi

# data/parameters/paras (GeauxDock, LA-SiGMA Application License 1.0)
PARAS_WEI = [0.999999999, -0.141543803, 0.075238177, 0.184487157, 0.308096607,
             0.092951333, -0.389581962, 0.171818907, 0.015012049]
PARAS_NOA = [0.775910, 18.970033, 0.005070, 0.532536, 2.046236, 0.525192, 110.815651,
             0.877866, 46.406089]
PARAS_NOB = [1.236367, -0.021163, -0.410825, 0.066766, 1.000000, -1.040155, -1.000000,
             -0.167652, -1.000000]
PARAS_PEL = (2.0, 0.5)
PARAS_PLJ = (0.881489, 0.742859, 0.700923)
PARAS_KDE = 0.5
PI_F = np.float32(3.1415926535)


def make_enepara(seed, identity_norm=False, weights=None):
    """Synthetic energy parameters (EnePara after OptimizeEnepara)."""
    rng = np.random.default_rng([seed, 0xE0E0])
    f32 = np.float32
    ep = abi.EnePara()
    sig = rng.uniform(3.0, 5.5, size=(abi.MAXTP1, abi.MAXTP2)).astype(f32)  # [prt][lig]
    eps = rng.uniform(0.1, 1.6, size=(abi.MAXTP1, abi.MAXTP2)).astype(f32)
    pmf0 = np.clip(sig + rng.normal(0, 0.3, size=sig.shape), 2.69, 7.38).astype(f32)
    pmf1 = np.clip(rng.normal(0.06, 1.0, size=sig.shape), -2.6, 6.5).astype(f32)
    has_hb = rng.random(size=sig.shape) < 0.13
    hdb0 = np.where(has_hb, rng.uniform(3.29, 4.55, size=sig.shape), 0.0).astype(f32)
    hdbs = np.where(has_hb, rng.uniform(0.6, 1.48, size=sig.shape), 1.0).astype(f32)
    lj = [f32(v) for v in PARAS_PLJ]
    for i in range(abi.MAXTP2):       # lig
        for j in range(abi.MAXTP1):   # prt
            tmp = f32(sig[j, i] * lj[2])
            ep.p1a[i][j] = f32(2.0) * eps[j, i] * f32(math.pow(float(tmp), 9.0))
            ep.p2a[i][j] = f32(3.0) * eps[j, i] * f32(math.pow(float(tmp), 6.0))
            ep.pmf0[i][j] = pmf0[j, i]
            ep.pmf1[i][j] = pmf1[j, i]
            ep.hdb0[i][j] = hdb0[j, i]
            ep.hdb1[i][j] = f32(1.0) / hdbs[j, i]
    ep.lj0, ep.lj1 = lj[0], lj[1]
    el0, el1 = f32(PARAS_PEL[0]), f32(PARAS_PEL[1])
    ep.el0, ep.el1 = el0, el1
    ep.a1 = f32(4.0) - f32(3.0) * el0
    ep.b1 = f32(2.0) * el0 - f32(3.0)
    hpp = rng.uniform(2.95, 9.03, size=abi.MAXTP4).astype(f32)
    lo, hi = hpp.min(), hpp.max()
    hpp = np.clip(((hpp - lo) / (hi - lo)) * 2.0 - 1.0, -1.0, 1.0).astype(f32)
    for i in range(abi.MAXTP4):
        ep.hpp[i] = hpp[i]
    sqrt_2_pi = f32(math.sqrt(float(f32(2.0) * PI_F)))
    for i in range(abi.MAXTP2):
        ep.hpl0[i] = f32(rng.uniform(-0.72, 0.29))
        ep.hpl1[i] = f32(rng.uniform(0.33, 0.64))
        ep.hpl2[i] = f32(math.log(float(f32(1.0) / (f32(ep.hpl1[i]) * sqrt_2_pi))))
    kde = f32(PARAS_KDE)
    ep.kde2 = f32(-0.5) / (kde * kde)
    ep.kde3 = f32(math.pow(float(kde * sqrt_2_pi), 3.0))
    w = weights if weights is not None else PARAS_WEI + [0.0]
    for i in range(abi.MAXWEI):
        ep.w[i] = f32(w[i])
        if identity_norm:
            ep.a_para[i], ep.b_para[i] = 1.0, 0.0
        else:
            ep.a_para[i] = f32(PARAS_NOA[i]) if i < 9 else 0.0
            ep.b_para[i] = f32(PARAS_NOB[i]) if i < 9 else 0.0
    return ep


def temperature_ladder(n_tmp, floor_t, ceiling_t):
    """SetTemperature (src/util.C:797-837): geometric ladder, k_B = 1."""
    f32 = np.float32
    t = np.zeros(n_tmp, np.float32)
    mb = np.zeros(n_tmp, np.float32)
    if n_tmp == 1:
        t[0] = floor_t
        mb[0] = f32(0.0) - f32(1.0) / f32(floor_t)
        return t, mb
    ratio = f32(math.exp(math.log(float(f32(ceiling_t) / f32(floor_t))) / float(n_tmp - 1)))
    a = f32(floor_t)
    for i in range(n_tmp):
        t[i] = a
        mb[i] = f32(0.0) - f32(1.0) / (f32(1.0) * a)
        a = f32(a * ratio)
    return t, mb


def _chain(rng, n, bond=1.5, angle_deg=109.5, jitter_deg=10.0, dihedrals=None):
    """Random bonded chain from internal coordinates (NeRF placement)."""
    ang = np.deg2rad(angle_deg + rng.uniform(-jitter_deg, jitter_deg, size=n))
    dih = dihedrals if dihedrals is not None else rng.uniform(-np.pi, np.pi, size=n)
    xyz = np.zeros((n, 3))
    if n > 1:
        xyz[1] = [bond, 0, 0]
    if n > 2:
        a = ang[2]
        xyz[2] = xyz[1] + bond * np.array([-math.cos(a), math.sin(a), 0.0])
    pts = [list(map(float, xyz[i])) for i in range(min(n, 3))]

    def cross(u, v):
        return [u[1] * v[2] - u[2] * v[1], u[2] * v[0] - u[0] * v[2], u[0] * v[1] - u[1] * v[0]]

    def unit(u):
        r = math.sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2])
        return [u[0] / r, u[1] / r, u[2] / r]

    for i in range(3, n):
        a, b, c = pts[i - 3], pts[i - 2], pts[i - 1]
        bc = unit([c[0] - b[0], c[1] - b[1], c[2] - b[2]])
        nrm = unit(cross([b[0] - a[0], b[1] - a[1], b[2] - a[2]], bc))
        m1 = cross(nrm, bc)
        th, ph = float(ang[i]), float(dih[i])
        d0 = -bond * math.cos(th)
        d1 = bond * math.sin(th) * math.cos(ph)
        d2 = bond * math.sin(th) * math.sin(ph)
        pts.append([c[k] + bc[k] * d0 + m1[k] * d1 + nrm[k] * d2 for k in range(3)])
    xyz[:] = np.array(pts[:n])
    return xyz, dih


class Complex:
    """Arrays of one complex, laid out as gd_complex_desc expects."""

    def desc(self):
        return abi.make_desc(self)

    @property
    def n_rep(self):
        return self.n_lig * self.n_prt * self.n_tmp


def make_complex(seed, lna=32, pnp=1200, pnk=100, pos=4, mcs_atoms=20, n_lig=1, n_prt=1,
                 n_tmp=8, box=None, void_radius=10.0, prt_sigma=0.5, t_floor=0.5,
                 t_ceiling=5.0, move_kind=abi.MOVE_GAUSSIAN, trans_sigma=0.2,
                 rot_sigma_deg=3.0, identity_norm=False, weights=None, enepara_seed=None):
    rng = np.random.default_rng([seed, 0xC0DE])
    f32 = np.float32
    cx = Complex()
    cx.seed = seed
    cx.n_lig, cx.n_prt, cx.n_tmp = n_lig, n_prt, n_tmp
    cx.lna, cx.pnp, cx.pnk, cx.pos = lna, pnp, pnk, pos
    if box is None:
        # keep the point density of 1,200 points in a 30 A cube with the void
        density = 1200.0 / (30.0 ** 3 - 4.0 / 3.0 * math.pi * void_radius ** 3)
        box = (pnp / density + 4.0 / 3.0 * math.pi * void_radius ** 3) ** (1.0 / 3.0)
    cx.box = box
    center = rng.uniform(-20.0, 20.0, size=3).astype(f32)
    cx.pocket_center = center

    # receptor
    pts = []
    while len(pts) < pnp:
        c = rng.uniform(-box / 2, box / 2, size=(2 * pnp, 3))
        c = c[np.linalg.norm(c, axis=1) >= void_radius]
        pts.extend(c.tolist())
    base = np.array(pts[:pnp]) + center
    prt = np.empty((n_prt, pnp, 3))
    prt[0] = base
    for k in range(1, n_prt):
        prt[k] = base + rng.normal(0.0, prt_sigma, size=base.shape)
    prt = prt.astype(f32)
    cx.prt_x = np.ascontiguousarray(prt[:, :, 0].reshape(-1))
    cx.prt_y = np.ascontiguousarray(prt[:, :, 1].reshape(-1))
    cx.prt_z = np.ascontiguousarray(prt[:, :, 2].reshape(-1))
    cx.prt_t = rng.integers(0, abi.MAXTP2, size=pnp).astype(np.int32)
    cx.prt_c = rng.integers(0, 4, size=pnp).astype(np.int32)
    prt_d = rng.integers(0, abi.MAXTP4, size=pnp).astype(np.int32)
    cx.prt_ele = np.clip(rng.normal(0.0, 0.3, size=pnp), -1.0, 1.0).astype(f32)
    cx.prt_seq3r = (np.arange(pnp) // 3).astype(np.int32)
    cx.n_res = int(cx.prt_seq3r.max()) + 1
    cx.prt_c0_and_d12_or_c2 = (((cx.prt_c == 0) & (prt_d == 12)) | (cx.prt_c == 2)).astype(np.int32)

    ep = make_enepara(seed if enepara_seed is None else enepara_seed,
                      identity_norm=identity_norm, weights=weights)
    hpp_tab = np.array(ep.hpp[:], dtype=f32)
    cx.prt_hpp = hpp_tab[prt_d].astype(f32)
    cx.enepara = np.frombuffer(bytes(ep), dtype=np.uint8).copy()
    cx.enepara_struct = ep

    # PSP [lig type][residue]
    psp = np.where(rng.random((abi.MAXTP2, cx.n_res)) < 0.1,
                   rng.normal(0.0, 1.0, (abi.MAXTP2, cx.n_res)), 0.0)
    cx.psp = np.ascontiguousarray(psp.astype(f32).reshape(-1))

    # ligand conformers
    best, best_r, dih0 = None, 1e9, None
    for _ in range(32):
        xyz, dih = _chain(rng, lna)
        xyz = xyz - xyz.mean(axis=0)
        r = np.linalg.norm(xyz, axis=1).max()
        if r < best_r:
            best, best_r, dih0 = xyz, r, dih
    confs = [best]
    for _ in range(1, n_lig):
        bestc, bestc_r = None, 1e9
        for _ in range(8):
            xyz, _d = _chain(rng, lna)
            xyz = xyz - xyz.mean(axis=0)
            r = np.linalg.norm(xyz, axis=1).max()
            if r < bestc_r:
                bestc, bestc_r = xyz, r
        confs.append(bestc)
    lig = np.array(confs).astype(f32)  # [n_lig][lna][3]
    cx.lig_x = np.ascontiguousarray(lig[:, :, 0].reshape(-1))
    cx.lig_y = np.ascontiguousarray(lig[:, :, 1].reshape(-1))
    cx.lig_z = np.ascontiguousarray(lig[:, :, 2].reshape(-1))
    cx.lig_t = rng.integers(0, abi.MAXTP2, size=lna).astype(np.int32)
    cx.lig_c = rng.normal(0.0, 0.2, size=lna).astype(f32)
    cx.lig_n = np.arange(lna, dtype=np.int32)

    # KDE points
    cx.kde_x, cx.kde_y, cx.kde_z = (
        np.ascontiguousarray((center[i] + rng.normal(0.0, 3.0, size=pnk)).astype(f32))
        for i in range(3))
    cx.kde_t = cx.lig_t[rng.integers(0, lna, size=pnk)].astype(np.int32)

    # MCS templates, indexed by atom number + 1
    cx.n_mcs = lna + 1
    start = lig[0] + center
    mx = np.full((pos, cx.n_mcs), abi.MCS_INVALID_COORD, f32)
    my, mz = mx.copy(), mx.copy()
    for m in range(pos):
        sel = rng.choice(lna, size=min(mcs_atoms, lna), replace=False)
        noise = rng.normal(0.0, 1.0, size=(len(sel), 3))
        for j, a in enumerate(sel):
            mx[m, a + 1] = start[a, 0] + noise[j, 0]
            my[m, a + 1] = start[a, 1] + noise[j, 1]
            mz[m, a + 1] = start[a, 2] + noise[j, 2]
    cx.mcs_x, cx.mcs_y, cx.mcs_z = mx.reshape(-1), my.reshape(-1), mz.reshape(-1)
    cx.mcs_tcc = (1.0 - rng.random(pos)).astype(f32)

    cx.temp_t, cx.temp_minus_beta = temperature_ladder(n_tmp, t_floor, t_ceiling)
    rot = math.radians(rot_sigma_deg)
    if move_kind == abi.MOVE_GAUSSIAN:
        cx.move_scale = np.array([trans_sigma] * 3 + [rot] * 3, f32)
    else:
        cx.move_scale = np.array([trans_sigma] * 3 + [rot] * 3, f32)
    cx.move_kind = move_kind
    return cx


# ---- the benchmark workloads ------------------------------------------------

def single_complex(seed=0, **kw):
    """One complex, CPU-runnable: 1,200-point receptor in a 30 A cube, 32-atom
    ligand, 1 x 1 x 8 replicas, KDE 100, MCS 4 x 20."""
    args = dict(lna=32, pnp=1200, pnk=100, pos=4, mcs_atoms=20, n_lig=1, n_prt=1, n_tmp=8,
                box=30.0)
    args.update(kw)
    return make_complex(seed, **args)


def ensemble(seed=1, **kw):
    """Ensemble: 3,000-point receptor in a 40 A cube with 11 conformers, 48-atom
    ligand with 20 conformers, 16 temperatures (3,520 replicas), KDE 200,
    MCS 8 x 40."""
    args = dict(lna=48, pnp=3000, pnk=200, pos=8, mcs_atoms=40, n_lig=20, n_prt=11,
                n_tmp=16, box=40.0)
    args.update(kw)
    return make_complex(seed, **args)


def uneven_batch(n_complex, seed=2, lna_range=(12, 64), pnp_range=(800, 4000), n_lig=5,
                 n_prt=3, n_tmp=8, **kw):
    """Uneven batch: ligand atoms U[12,64], receptor points U[800,4000], each
    complex 5 ligand x 3 protein conformers x 8 temperatures."""
    rng = np.random.default_rng([seed, 0xBA7C])
    jobs = []
    for i in range(n_complex):
        lna = int(rng.integers(lna_range[0], lna_range[1] + 1))
        pnp = int(rng.integers(pnp_range[0], pnp_range[1] + 1))
        args = dict(lna=lna, pnp=pnp, pnk=100, pos=4, mcs_atoms=min(20, lna), n_lig=n_lig,
                    n_prt=n_prt, n_tmp=n_tmp, enepara_seed=seed)
        args.update(kw)
        jobs.append((seed * 100003 + i, args))
    return _build_all(jobs)


def _build_one(job):
    return make_complex(job[0], **job[1])


def _gpu_touched():
    """True once this process has loaded the engine library or initialized
    torch's GPU context: forking is unsafe from then on."""
    import sys
    pkg = sys.modules.get(__package__)
    if pkg is not None and getattr(pkg, "_lib", None) is not None:
        return True
    torch = sys.modules.get("torch")
    return bool(torch is not None and torch.cuda.is_initialized())


def _build_all(jobs, min_parallel=64):
    """make_complex over (seed, kwargs) jobs, in order. Every complex depends
    only on its own seed, so large batches are built by a forked process pool
    while the process has not touched the GPU (sequentially afterwards)."""
    if len(jobs) < min_parallel or _gpu_touched():
        return [_build_one(j) for j in jobs]
    import concurrent.futures as cf
    import multiprocessing as mp
    workers = max(1, min(16, (os.cpu_count() or 1)))
    with cf.ProcessPoolExecutor(workers, mp_context=mp.get_context("fork")) as ex:
        out = list(ex.map(_build_one, jobs, chunksize=max(1, len(jobs) // (4 * workers))))
    return out
