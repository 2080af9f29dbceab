"""Pins the CPU oracle (oracle/gdo.c) before it is trusted as the checker.

Each check is chosen so a plausible mistake in the oracle fails it: published
known-answer vectors (Philox), closed forms and limits of each energy term
(src/kernel_cuda_l2_calcenergy.cu), invariants (rotation orthonormality,
exchange permutations, detailed balance), and brute force on tiny complexes
with asymmetric tables (catches transposed [lig][prt] / [prt][lig] indexing).
"""
import ctypes as C
import math

import numpy as np
import pytest

from geaxdock_gpu_amd import abi, synth


# ---------------------------------------------------------------- random numbers

# Random123 kat_vectors, philox4x32_10
KAT = [
    ([0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
    ([0xffffffff] * 4, [0xffffffff] * 2, [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
    ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
     [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]),
]


@pytest.mark.parametrize("ctr,key,want", KAT)
def test_philox_known_answers(oracle, ctr, key, want):
    assert oracle.philox(ctr, key).tolist() == want


def test_uniform_range(oracle):
    lib = oracle.lib()
    assert lib.gdo_uniform(0) == 2.0 ** -24
    assert lib.gdo_uniform(0xFFFFFFFF) == 1.0
    assert lib.gdo_uniform(0x80000000) == (0x800000 + 1) * 2.0 ** -24


def test_gaussian_moves_have_unit_variance(oracle):
    lib = oracle.lib()
    scale = np.ones(6, np.float32)
    u = np.zeros(8, np.float32)
    dm = np.zeros(6, np.float32)
    fp = C.POINTER(C.c_float)
    zs = []
    for s in range(20000):
        lib.gdo_mc_uniforms(7, 0, 3, s, u.ctypes.data_as(fp))
        lib.gdo_move_offsets(u.ctypes.data_as(fp), scale.ctypes.data_as(fp), abi.MOVE_GAUSSIAN,
                             dm.ctypes.data_as(fp))
        zs.append(dm.copy())
    z = np.array(zs)
    assert np.all(np.abs(z.mean(0)) < 0.03)
    assert np.all(np.abs(z.std(0) - 1.0) < 0.03)
    assert abs(np.corrcoef(z[:, 0], z[:, 1])[0, 1]) < 0.03


# ------------------------------------------------------------------- decision exp

def test_expf_accuracy_and_edges(oracle):
    xs = np.concatenate([np.linspace(-87, 88, 200001, dtype=np.float32),
                         np.float32([0.0, -0.0, 1e-30, -1e-30, 0.5, -0.5])])
    got = np.array([oracle.expf(x) for x in xs], np.float64)
    want = np.exp(xs.astype(np.float64))
    ulp = np.spacing(want.astype(np.float32)).astype(np.float64)
    assert np.max(np.abs(got - want) / ulp) <= 2.0
    assert oracle.expf(0.0) == 1.0
    assert math.isnan(oracle.expf(float("nan")))
    assert oracle.expf(88.5) == float("inf") and oracle.expf(-87.5) == 0.0
    assert np.all(np.diff(got[:200001]) >= 0)  # monotone


def test_accept_semantics(oracle):
    acc = oracle.lib().gdo_accept
    assert acc(1.0, 2.0, -1.0, 1.0 - 2 ** -24) == 1  # downhill, U < 1
    assert acc(1.0, 1.0, -1.0, 1.0) == 0             # strict U < exp(0)
    assert acc(float("nan"), 1.0, -1.0, 2 ** -24) == 0
    assert acc(float("inf"), 1.0, -1.0, 2 ** -24) == 0
    # far from the threshold the decision equals U < exp(dE * minus_beta) in FP64
    rng = np.random.default_rng(0)
    for _ in range(20000):
        de = rng.normal(0, 2)
        mb = -1.0 / rng.uniform(0.3, 5.0)
        u = (rng.integers(0, 2 ** 24) + 1) * 2.0 ** -24
        p = math.exp(np.float32(de) * np.float32(mb))
        if abs(u - p) > 1e-5:
            assert acc(float(np.float32(de)), 0.0, float(np.float32(mb)), u) == int(u < p)


def _exchange(oracle, n_tmp, mode_t, et, mb, u, it, ac):
    et = np.ascontiguousarray(et, np.float32)
    mb = np.ascontiguousarray(mb, np.float32)
    u = np.ascontiguousarray(u, np.float32)
    it = np.array(it, np.int32)
    ac = np.array(ac, np.int32)
    fp = C.POINTER(C.c_float)
    ip = C.POINTER(C.c_int32)
    oracle.lib().gdo_exchange_group(n_tmp, mode_t, et.ctypes.data_as(fp), mb.ctypes.data_as(fp),
                                    u.ctypes.data_as(fp), it.ctypes.data_as(ip),
                                    ac.ctypes.data_as(ip))
    return it, ac


def test_exchange_pairs_attempted(oracle):
    """Mode 0 pairs temperatures (0,1),(2,3),...; mode 1 pairs (1,2),(3,4),...
    (src/kernel_cuda_l1_exchangereplicas.cu:43-47). With U tiny every attempted
    pair swaps, so the result shows exactly which pairs were attempted."""
    _, mb = synth.temperature_ladder(8, 0.5, 5.0)
    et = np.zeros(8)
    u = np.full(4, 2 ** -24)
    it, ac = _exchange(oracle, 8, 0, et, mb, u, np.arange(8), np.zeros(8))
    assert it.tolist() == [1, 0, 3, 2, 5, 4, 7, 6] and ac.tolist() == [1] * 8
    it, ac = _exchange(oracle, 8, 1, et, mb, u, np.arange(8), np.zeros(8))
    assert it.tolist() == [0, 2, 1, 4, 3, 6, 5, 7] and ac.tolist() == [0, 1, 1, 1, 1, 1, 1, 0]
    _, mb7 = synth.temperature_ladder(7, 0.5, 5.0)
    it, _ = _exchange(oracle, 7, 1, np.zeros(7), mb7, np.full(3, 2 ** -24), np.arange(7), np.zeros(7))
    assert it.tolist() == [0, 2, 1, 4, 3, 6, 5]


def test_exchange_direction_and_detailed_balance(oracle):
    """A cold replica holding the higher energy always swaps with its hot
    neighbour; the reverse swap happens with probability exp(-(b1-b2)(E1-E2)),
    which is the ratio of Boltzmann weights of the two assignments."""
    mb = np.float32([-2.0, -0.5])
    it, _ = _exchange(oracle, 2, 0, [5.0, 1.0], mb, [1.0], [0, 1], [0, 0])
    assert it.tolist() == [1, 0]
    p = math.exp((-2.0 + 0.5) * (5.0 - 1.0))  # cold slot has the lower energy
    rng = np.random.default_rng(3)
    n, sw = 40000, 0
    for _ in range(n):
        u = (rng.integers(0, 2 ** 24) + 1) * 2.0 ** -24
        it, _ = _exchange(oracle, 2, 0, [1.0, 5.0], mb, [u], [0, 1], [0, 0])
        sw += it[0] == 1
    assert abs(sw / n - p) < 4 * math.sqrt(p * (1 - p) / n)


def test_exchange_is_permutation(oracle):
    rng = np.random.default_rng(5)
    for n_tmp in (2, 3, 8, 16, 40):
        _, mb = synth.temperature_ladder(n_tmp, 0.5, 5.0)
        for _ in range(50):
            it0 = rng.permutation(n_tmp)
            it, ac = _exchange(oracle, n_tmp, int(rng.integers(0, 2)), rng.normal(0, 3, n_tmp), mb,
                               rng.uniform(0, 1, n_tmp // 2), it0, np.zeros(n_tmp))
            assert sorted(it.tolist()) == list(range(n_tmp))
            assert ac.sum() == 2 * np.sum(it != it0) // 2 * 1 or ac.sum() % 2 == 0


# ------------------------------------------------------------- tiny complexes

def tiny_complex(lig_xyz, prt_xyz, lig_t=None, prt_t=None, prt_c=None, flag=None, q=None, ele=None,
                 hpp=None, kde=None, mcs=None, pocket=(0.0, 0.0, 0.0), ep=None, n_res=4, psp=None):
    """A complex built directly from coordinates (ligand frame = pocket frame)."""
    cx = synth.Complex()
    lig = np.asarray(lig_xyz, np.float32).reshape(-1, 3)
    prt = np.asarray(prt_xyz, np.float32).reshape(-1, 3)
    lna, pnp = len(lig), len(prt)
    cx.n_lig = cx.n_prt = cx.n_tmp = 1
    cx.lna, cx.pnp = lna, pnp
    cx.pocket_center = np.asarray(pocket, np.float32)
    cx.lig_x, cx.lig_y, cx.lig_z = (np.ascontiguousarray(lig[:, i]) for i in range(3))
    cx.lig_t = np.asarray(lig_t if lig_t is not None else [0] * lna, np.int32)
    cx.lig_c = np.asarray(q if q is not None else [0.0] * lna, np.float32)
    cx.lig_n = np.arange(lna, dtype=np.int32)
    cx.prt_x, cx.prt_y, cx.prt_z = (np.ascontiguousarray(prt[:, i] + pocket[i]) for i in range(3))
    cx.prt_t = np.asarray(prt_t if prt_t is not None else [0] * pnp, np.int32)
    cx.prt_c = np.asarray(prt_c if prt_c is not None else [1] * pnp, np.int32)
    cx.prt_c0_and_d12_or_c2 = np.asarray(flag if flag is not None else [0] * pnp, np.int32)
    cx.prt_ele = np.asarray(ele if ele is not None else [0.0] * pnp, np.float32)
    cx.prt_hpp = np.asarray(hpp if hpp is not None else [0.0] * pnp, np.float32)
    cx.prt_seq3r = np.arange(pnp, dtype=np.int32) % n_res
    cx.n_res = n_res
    cx.psp = np.ascontiguousarray(psp if psp is not None else np.zeros(24 * n_res), np.float32)
    kde = kde if kde is not None else []
    cx.pnk = len(kde)
    k = np.asarray([p[:3] for p in kde], np.float32).reshape(-1, 3)
    cx.kde_x, cx.kde_y, cx.kde_z = (np.ascontiguousarray(k[:, i] + pocket[i]) for i in range(3))
    cx.kde_t = np.asarray([p[3] for p in kde], np.int32)
    mcs = mcs if mcs is not None else []
    cx.pos = len(mcs)
    cx.n_mcs = lna + 1
    mx = np.full((max(cx.pos, 1), lna + 1), abi.MCS_INVALID_COORD, np.float32)
    my, mz = mx.copy(), mx.copy()
    tcc = []
    for m, (t, pts) in enumerate(mcs):
        tcc.append(t)
        for a, xyz in pts.items():
            mx[m, a + 1], my[m, a + 1], mz[m, a + 1] = (xyz[i] + pocket[i] for i in range(3))
    cx.mcs_x, cx.mcs_y, cx.mcs_z = mx.reshape(-1), my.reshape(-1), mz.reshape(-1)
    cx.mcs_tcc = np.asarray(tcc if tcc else [1.0], np.float32)
    ep = ep if ep is not None else synth.make_enepara(0, identity_norm=True,
                                                      weights=[1.0] * 9 + [0.0])
    cx.enepara = np.frombuffer(bytes(ep), dtype=np.uint8).copy()
    cx.enepara_struct = ep
    cx.temp_t, cx.temp_minus_beta = synth.temperature_ladder(1, 1.0, 1.0)
    cx.move_scale = np.zeros(6, np.float32)
    cx.move_kind = abi.MOVE_GAUSSIAN
    return cx


def e64(oracle, cx, mm=None, ref=True):
    d = cx.desc()
    r = oracle.refmap_f64(d, cx) if ref else None
    return oracle.energy_f64(d, r, 0, 0, np.zeros(6, np.float32) if mm is None else mm)


def test_vdw_limits_and_root(oracle):
    cx0 = tiny_complex([[0, 0, 0]], [[3, 0, 0]])
    ep = cx0.enepara_struct
    p1a, p2a, lj0, lj1 = ep.p1a[0][0], ep.p2a[0][0], ep.lj0, ep.lj1
    # root where p1a/d^9 = p2a/d^6
    d0 = (p1a / p2a) ** (1.0 / 3.0)
    lo = e64(oracle, tiny_complex([[0, 0, 0]], [[d0 * 0.98, 0, 0]])).raw[0]
    hi = e64(oracle, tiny_complex([[0, 0, 0]], [[d0 * 1.02, 0, 0]])).raw[0]
    assert lo > 0 > hi
    # near overlap the soft core tends to 1 / (lj0 (1 + lj1 d^2))
    d = 1e-3
    v = e64(oracle, tiny_complex([[0, 0, 0]], [[d, 0, 0]])).raw[0]
    assert v == pytest.approx(1.0 / (lj0 * (1 + lj1 * d * d)), rel=1e-6)
    # far tail is the attractive -p2a / d^6 (relative correction p1/p2 ~ d^-3)
    d = 400.0
    v = e64(oracle, tiny_complex([[0, 0, 0]], [[d, 0, 0]])).raw[0]
    assert v == pytest.approx(-p2a / d ** 6, rel=1e-6)


def test_ele_switch_continuity_and_sign(oracle):
    ep = synth.make_enepara(0, identity_norm=True)
    el1 = ep.el1
    s_at = 1.0 / el1
    f = lambda d, q: e64(oracle, tiny_complex([[0, 0, 0]], [[d, 0, 0]], q=[q], ele=[1.0])).raw[1]
    below, above = f(s_at * (1 - 1e-7), 0.5), f(s_at * (1 + 1e-7), 0.5)
    assert below == pytest.approx(above, rel=1e-5)
    assert f(s_at, 0.5) == pytest.approx(0.5, rel=1e-6)  # g(1) = 1
    h = 1e-4 * s_at  # g'(1) = -1 in s: slope in d is -el1 * q
    slope = (f(s_at * (1 + 1e-3) + h, 0.5) - f(s_at * (1 + 1e-3) - h, 0.5)) / (2 * h)
    assert slope == pytest.approx(-0.5 * el1 / (1 + 1e-3) ** 2, rel=1e-3)
    assert f(7.0, -0.5) == pytest.approx(-f(7.0, 0.5))
    assert f(7.0, 0.5) == pytest.approx(0.5 / (el1 * 7.0), rel=1e-6)


def test_pmf_midpoints_and_limits(oracle):
    cx = tiny_complex([[0, 0, 0]], [[3, 0, 0]])
    ep = cx.enepara_struct
    pmf0, pmf1 = ep.pmf0[0][0], ep.pmf1[0][0]
    f = lambda d: e64(oracle, tiny_complex([[0, 0, 0]], [[d, 0, 0]])).raw[2]
    assert f(pmf0) == pytest.approx(pmf1 / 2, rel=1e-6)
    assert f(12.0) == pytest.approx(pmf1 / 2, rel=1e-6)  # (6 - d/2) vanishes at 12 A
    assert f(30.0) == pytest.approx(pmf1, rel=1e-9)
    assert f(0.05) == pytest.approx(pmf1, rel=1e-4)
    assert abs(f(0.5 * (pmf0 + 12.0))) < abs(pmf1) * 1e-2


def test_hdb_peak_symmetry_and_switch(oracle):
    ep = synth.make_enepara(0, identity_norm=True)
    pairs = [(lt, pt) for lt in range(24) for pt in range(24) if ep.hdb0[lt][pt] > 0.1]
    lt, pt = pairs[0]
    h0, h1 = ep.hdb0[lt][pt], ep.hdb1[lt][pt]
    f = lambda d, pt=pt: e64(oracle, tiny_complex([[0, 0, 0]], [[d, 0, 0]], lig_t=[lt],
                                                  prt_t=[pt])).raw[4]
    assert f(h0) == pytest.approx(-h1 / math.sqrt(2 * math.pi), rel=1e-6)
    assert f(h0 - 0.7) == pytest.approx(f(h0 + 0.7), rel=1e-5)
    off = [p for p in range(24) if ep.hdb0[lt][p] <= 0.1][0]
    assert f(h0, pt=off) == 0.0


def test_hpc_switch_and_restraint(oracle):
    ep = synth.make_enepara(0, identity_norm=True)
    hpl0, hpl1, hpl2 = ep.hpl0[0], ep.hpl1[0], ep.hpl2[0]

    def f(d, flag=1):
        return e64(oracle, tiny_complex([[0, 0, 0]], [[d, 0, 0]], flag=[flag], hpp=[0.8])).raw[5]
    closed = lambda s: 0.5 * ((0.8 * s - hpl0) / hpl1) ** 2 - hpl2
    assert f(1e-4) == pytest.approx(closed(1.0), rel=1e-6)   # switch = 1 at d = 0
    assert f(9.0) == pytest.approx(closed(0.0), rel=1e-6)    # switch = 0 at d = 9
    assert f(9.5) == pytest.approx(closed(0.0), rel=1e-9)
    assert f(2.0, flag=0) == pytest.approx(closed(0.0), rel=1e-9)
    x = 4.5 ** 2 / 81
    assert f(4.5) == pytest.approx(closed(1 - (3.5 * x - 4.5 * x * x + 2.5 * x ** 3 - 0.5 * x ** 4)),
                                   rel=1e-6)


def test_psp_contact_only(oracle):
    ep = synth.make_enepara(0, identity_norm=True)
    pmf0 = ep.pmf0[3][5]
    psp = np.zeros(24 * 4, np.float32)
    psp[3 * 4 + 2] = 0.7  # psp[lig type 3][residue 2]
    mk = lambda d, c: tiny_complex([[0, 0, 0]], [[9, 9, 9], [9, -9, 9], [d, 0, 0]], lig_t=[3],
                                   prt_t=[5, 5, 5], prt_c=[1, 1, c], psp=psp)
    assert e64(oracle, mk(pmf0 * 0.99, 2)).raw[3] == pytest.approx(0.7, rel=1e-6)
    assert e64(oracle, mk(pmf0 * 1.01, 2)).raw[3] == 0.0
    assert e64(oracle, mk(pmf0 * 0.99, 1)).raw[3] == 0.0


def test_kde_single_point(oracle):
    ep = synth.make_enepara(0, identity_norm=True)
    bw = 0.5
    r = 0.4
    cx = tiny_complex([[0, 0, 0], [5, 0, 0]], [[20, 0, 0]], lig_t=[2, 4],
                      kde=[(0, r, 0, 2), (9, 9, 9, 7)])
    want = math.exp(-r * r / (2 * bw * bw)) / ((bw * math.sqrt(2 * math.pi)) ** 3) / 2
    assert e64(oracle, cx).raw[6] == pytest.approx(want, rel=1e-5)
    assert ep.kde3 == pytest.approx((bw * math.sqrt(2 * math.pi)) ** 3, rel=1e-6)


def test_mcs_uniform_offset(oracle):
    lig = [[0, 0, 0], [1.5, 0, 0], [1.5, 1.5, 0]]
    delta = np.array([0.3, -0.4, 1.2])
    pts = {a: np.array(lig[a]) + delta for a in range(3)}
    cx = tiny_complex(lig, [[20, 0, 0]], mcs=[(0.6, pts), (0.2, {0: np.array(lig[0]) + 2 * delta})])
    want = math.log((0.6 * np.linalg.norm(delta) + 0.2 * 2 * np.linalg.norm(delta)) / 2)
    assert e64(oracle, cx).raw[7] == pytest.approx(want, rel=1e-6)


def test_mcs_template_on_the_pose_gives_log_zero(oracle):
    """Every template atom exactly on the pose: the reference's log of the
    mean template distance is log(0) = -inf (src/kernel_cuda_l2_calcenergy.cu:341),
    kept as is; with a positive weight the total is -inf too."""
    lig = [[0, 0, 0], [1.5, 0, 0], [1.5, 1.5, 0]]
    pts = {a: np.array(lig[a]) for a in range(3)}
    cx = tiny_complex(lig, [[20, 0, 0]], mcs=[(0.6, pts), (0.2, {1: np.array(lig[1])})])
    o = e64(oracle, cx)
    assert o.raw[7] == -math.inf and o.e[7] == -math.inf and o.e[9] == -math.inf
    d = cx.desc()
    f = oracle.energy_f32(d, oracle.refmap_f32(d, cx), 0, 0, np.zeros(6, np.float32))
    assert f.e[7] == -math.inf and f.e[9] == -math.inf


def test_dst_rmsd_and_rigid_placement(oracle):
    rng = np.random.default_rng(2)
    lig = rng.normal(0, 2, (6, 3))
    lig -= lig.mean(0)
    cx = tiny_complex(lig, [[20, 0, 0]], pocket=(3.0, -4.0, 7.0))
    t = np.float32([0.5, -1.25, 2.0])
    mm = np.concatenate([t, np.zeros(3, np.float32)])
    o = e64(oracle, cx, mm)
    assert o.raw[8] == pytest.approx(np.linalg.norm(t), rel=1e-7)       # Q16: dst = |t|
    assert o.rmsd == pytest.approx(np.linalg.norm(t), rel=1e-6)         # pure translation
    d = cx.desc()
    mm = np.float32([0, 0, 0, 0.7, -1.1, 2.3])
    xyz = oracle.place_f64(d, cx, 0, mm) - np.float64(cx.pocket_center)[:, None]
    R, *_ = np.linalg.lstsq(lig.astype(np.float64), xyz.T, rcond=None)
    assert np.allclose(R.T @ R, np.eye(3), atol=1e-6)
    assert np.linalg.det(R) == pytest.approx(1.0, abs=1e-6)
    # the matrix of src/kernel_cuda_l2_move.cu:43-51 is Rz(a) Ry(b) Rx(g)
    a, b, g = (float(v) for v in mm[3:])
    Rz = np.array([[math.cos(a), -math.sin(a), 0], [math.sin(a), math.cos(a), 0], [0, 0, 1]])
    Ry = np.array([[math.cos(b), 0, math.sin(b)], [0, 1, 0], [-math.sin(b), 0, math.cos(b)]])
    Rx = np.array([[1, 0, 0], [0, math.cos(g), -math.sin(g)], [0, math.sin(g), math.cos(g)]])
    assert np.allclose(R.T, Rz @ Ry @ Rx, atol=1e-6)
    assert np.allclose(oracle.place_f64(d, cx, 0, np.zeros(6, np.float32)),
                       (lig + np.float64(cx.pocket_center)).T, atol=1e-6)


def test_unopt_weights_reduce_to_vdw_plus_dst(oracle):
    """With WEI 1 0 ... 0 1, NOA 1, NOB 0 (data/parameters/unopt_paras) the
    total is vdw/lna + |t|."""
    cx = synth.single_complex(seed=5, identity_norm=True, weights=[1.0] + [0.0] * 7 + [1.0, 0.0])
    t = np.float32([0.3, 0.2, -0.4, 0.1, 0.2, 0.3])
    o = e64(oracle, cx, t)
    assert o.e[9] == pytest.approx(o.raw[0] + np.linalg.norm(t[:3].astype(np.float64)), rel=1e-12)


def test_mcc_start_pose_and_empty_map(oracle):
    cx = synth.single_complex(seed=4, void_radius=5.5)
    o = e64(oracle, cx)
    assert o.tp > 0 and o.tn > 0 and o.fp == 0 and o.fn == 0 and o.cms == 1.0
    cx = synth.single_complex(seed=4, void_radius=12.0)
    o = e64(oracle, cx)
    assert o.tp == 0 and o.cms == 0.0  # empty reference map: denominator 0 -> 0


def test_brute_force_tiny_asymmetric_tables(oracle):
    """Independent double loops over a 3-atom x 5-point complex whose type
    tables differ under transposition; indexes follow EnePara's
    [ligand type][receptor type] and Psp's [ligand type][residue]."""
    rng = np.random.default_rng(9)
    lig = rng.normal(0, 1.2, (3, 3))
    prt = rng.normal(0, 3.0, (5, 3)) + np.array([0, 0, 4.5])
    prt[0] = lig[0] + [2.0, 0.5, 0.0]   # class-2 contacts so the PSP term is exercised
    prt[2] = lig[1] + [0.0, 2.5, 0.3]
    lt, pt = [1, 5, 7], [2, 9, 7, 5, 1]
    pc, fl = [2, 0, 2, 1, 3], [1, 1, 1, 0, 0]
    q, ele, hpp = rng.normal(0, .3, 3), rng.normal(0, .3, 5), rng.uniform(-1, 1, 5)
    psp = rng.normal(0, 1, 24 * 4).astype(np.float32)
    ep = synth.make_enepara(4, identity_norm=True, weights=[1.0] * 9 + [0.0])
    for i in range(24):
        for j in range(30):
            ep.pmf0[i][j] = 4.0 + 0.05 * i - 0.03 * j
            ep.hdb0[i][j] = 3.5 + 0.01 * i if (i + j) % 3 == 0 else 0.0
    cx = tiny_complex(lig, prt, lig_t=lt, prt_t=pt, prt_c=pc, flag=fl, q=q, ele=ele, hpp=hpp, ep=ep,
                      psp=psp)
    o = e64(oracle, cx)
    L = lambda a: np.float64(np.float32(a))
    ev = ee = epm = eps = eh = ehp = 0.0
    for l in range(3):
        hs = 0.0
        for p in range(5):
            d = np.linalg.norm(L(lig[l]) - L(prt[p]))
            a, b = lt[l], pt[p]
            p1, p2 = ep.p1a[a][b] / d ** 9, ep.p2a[a][b] / d ** 6
            ev += (p1 - p2) / (p1 * ep.lj0 * (1 + ep.lj1 * d * d) + 1)
            s = ep.el1 * d
            g = ep.el0 + ep.a1 * s * s + ep.b1 * s ** 3 if s < 1 else 1 / s
            ee += L(q[l]) * L(ele[p]) * g
            epm += ep.pmf1[a][b] / (1 + math.exp((-0.5 * d + 6) * (d - ep.pmf0[a][b])))
            if pc[p] == 2 and d <= ep.pmf0[a][b]:
                eps += psp[a * 4 + p % 4]
            if ep.hdb0[a][b] > 0.1:
                eh += -ep.hdb1[a][b] / math.sqrt(2 * math.pi) * math.exp(
                    -0.5 * ((d - ep.hdb0[a][b]) * ep.hdb1[a][b]) ** 2)
            if fl[p] == 1 and d * d <= 81:
                x = d * d / 81
                hs += L(hpp[p]) * (1 - (3.5 * x - 4.5 * x * x + 2.5 * x ** 3 - 0.5 * x ** 4))
        ehp += 0.5 * ((hs - ep.hpl0[lt[l]]) / ep.hpl1[lt[l]]) ** 2 - ep.hpl2[lt[l]]
    want = np.array([ev, ee, epm, eps, eh, ehp]) / 3
    np.testing.assert_allclose(np.array(o.raw[:6]), want, rtol=1e-7, atol=1e-12)
    assert eps != 0.0 and eh != 0.0 and ehp != 0.0


def test_fp32_cpu_path_vs_fp64(oracle):
    cx = synth.single_complex(seed=3, identity_norm=True, void_radius=7.0)
    d = cx.desc()
    r64, r32 = oracle.refmap_f64(d, cx), oracle.refmap_f32(d, cx)
    rng = np.random.default_rng(1)
    for i in range(12):
        mm = np.zeros(6, np.float32)
        if i:
            mm[:3] = rng.normal(0, 1.5, 3)
            mm[3:] = rng.uniform(-3, 3, 3)
        o = oracle.energy_f64(d, r64, 0, 0, mm)
        f = oracle.energy_f32(d, r32, 0, 0, mm)
        for t in range(10):
            assert abs(f.e[t] - o.e[t]) <= 1e-4 * max(o.mag[t], 1e-30), (i, t, f.e[t], o.e[t])


def test_temperature_ladder_closed_form():
    t, mb = synth.temperature_ladder(8, 0.5, 5.0)
    r = (5.0 / 0.5) ** (1 / 7)
    np.testing.assert_allclose(t, 0.5 * r ** np.arange(8), rtol=1e-5)
    np.testing.assert_allclose(mb, -1.0 / t, rtol=1e-6)
    assert t[-1] == pytest.approx(5.0, rel=1e-5)


def test_cpu_mc_path_deterministic_and_consistent(oracle):
    """The CPU path is a function of the seed; records hold the energy of the
    recorded pose; counters match the record counts."""
    cx = synth.single_complex(seed=1)
    d = cx.desc()
    runs = []
    for _ in range(2):
        st = oracle.McState(cx, rec_cap=121)
        oracle.mc_run(d, st, seed=3, init=True, n_steps=0, nc=20)
        oracle.mc_run(d, st, seed=3, n_steps=120, nc=20, threads=2)
        runs.append(st)
    a, b = runs
    assert np.array_equal(a.mm, b.mm) and np.array_equal(a.acs_x, b.acs_x)
    assert np.array_equal(a.rec_cnt, a.acs_mc + 1)
    ref = oracle.refmap_f32(d, cx)
    r = a.records(0)[-1]
    e = oracle.energy_f32(d, ref, 0, 0, r["movematrix"])
    assert np.float32(e.e[9]) == r["energy"]["e"][9]
    assert sorted(a.idx_tmp.tolist()) == list(range(8))


def test_da_chain_reduces_to_plain_when_surrogate_is_exact(oracle):
    """Pins gdo_mc_run_da: with every receptor-pair weight zero the exact
    energy equals the surrogate (the cheap terms), so every composite move is
    accepted and the delayed-acceptance chain must pass through exactly the
    plain chain's states every k steps (same uniforms per sub-step). A wrong
    sign or a dropped surrogate correction in the composite test, or wrong
    sub-step uniforms, breaks the equality."""
    w = [0.0] * 6 + [0.7, 0.4, 1.0, 0.0]
    cx = synth.make_complex(seed=12, lna=12, pnp=300, n_lig=4, n_prt=1, n_tmp=4, box=20.0,
                            void_radius=6.0, weights=w, t_floor=0.5, t_ceiling=5.0)
    d = cx.desc()
    k, n = 4, 160
    plain = oracle.McState(cx, rec_cap=n + 1)
    oracle.mc_run(d, plain, seed=9, init=True, step0=0, n_steps=0, nc=0)
    da = oracle.McState(cx, rec_cap=n + 1)
    oracle.mc_run(d, da, seed=9, init=True, step0=0, n_steps=0, nc=0)
    moved = 0
    for s0 in range(0, n, k):
        before = plain.mm.copy()
        oracle.mc_run(d, plain, seed=9, init=False, step0=s0, n_steps=k, nc=0)
        oracle.mc_run_da(d, da, seed=9, step0=s0, n_steps=k, nc=0, k_sub=k)
        assert np.array_equal(plain.mm, da.mm), s0
        assert np.array_equal(plain.e_old, da.e_old), s0
        moved += int(np.any(plain.mm != before, axis=1).sum())
    assert moved > n // k  # the chains did move
    # a chain with exchanges: composite groups align with exchange segments
    p2 = oracle.McState(cx, rec_cap=0)
    d2 = oracle.McState(cx, rec_cap=0)
    for st in (p2, d2):
        oracle.mc_run(d, st, seed=3, init=True, step0=0, n_steps=0, nc=20)
    oracle.mc_run(d, p2, seed=3, init=False, step0=0, n_steps=200, nc=20)
    oracle.mc_run_da(d, d2, seed=3, step0=0, n_steps=200, nc=20, k_sub=k)
    assert np.array_equal(p2.mm, d2.mm) and np.array_equal(p2.idx_tmp, d2.idx_tmp)


def _load_trajectory():
    import os
    from geaxdock_gpu_amd import io
    g = os.path.join(os.path.dirname(__file__), "golden", "inputs")
    cx = io.load_complex(os.path.join(g, "1a07C1", "1a07C1.sdf"),
                         os.path.join(g, "1a07C1", "1a07C.pdb"),
                         os.path.join(g, "1a07C1", "1a07C1-0.8.ff"),
                         os.path.join(g, "parameters", "paras"), complex_id="1a07C1")
    with open(os.path.join(g, "trajectory.out")) as f:
        rows = [l.split() for l in f.read().split("\n")[1:] if l.strip()]
    ip = np.array([int(r[0]) for r in rows])
    il = np.array([int(r[1]) for r in rows])
    mm = np.array([[float(v) for v in r[2:8]] for r in rows], np.float32)
    cms = np.array([float(r[8]) for r in rows])
    ener = np.array([float(r[9]) for r in rows])
    return cx, ip, il, mm, cms, ener


def test_reference_trajectory_known_answers(oracle):
    """data/trajectory.out: 2,666 poses of a reference run on 1a07C1 (receptor
    conformer, ligand conformer, move vector) with the MCC and total energy
    the reference printed (6 decimals). Loaded through the library's input
    loader, the FP64 oracle reproduces every total energy within 1e-5 and
    every MCC within 2e-6 (print rounding 5e-7); the FP32 CPU path within
    2e-5 and 2e-6."""
    cx, ip, il, mm, cms, ener = _load_trajectory()
    assert len(ener) == 2666
    d = cx.desc()
    r64, r32 = oracle.refmap_f64(d, cx), oracle.refmap_f32(d, cx)
    worst = np.zeros(4)
    for i in range(len(ener)):
        o = oracle.energy_f64(d, r64, int(ip[i]), int(il[i]), mm[i])
        f = oracle.energy_f32(d, r32, int(ip[i]), int(il[i]), mm[i])
        errs = [abs(o.e[9] - ener[i]), abs(o.cms - cms[i]), abs(f.e[9] - ener[i]),
                abs(f.cms - cms[i])]
        worst = np.maximum(worst, errs)
        assert errs[0] <= 1e-5 and errs[1] <= 2e-6, (i, o.e[9], ener[i], o.cms, cms[i])
        assert errs[2] <= 2e-5 and errs[3] <= 2e-6, (i, f.e[9], ener[i], f.cms, cms[i])
    print("\nworst |d ener| f64 %.2e, |d cms| f64 %.2e, |d ener| f32 %.2e, |d cms| f32 %.2e"
          % tuple(worst))
    # the file spans a real descent: a broken term could not match all rows
    assert ener.max() - ener.min() > 2.0
