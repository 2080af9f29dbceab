"""GPU parity of the engine against the CPU oracle, through the C ABI.

Energies: every term of gd_calc_energy (FP32, the MC kernel's device code)
against the FP64 oracle (oracle/gdo.c, restating
src/kernel_cuda_l2_calcenergy.cu). Tolerance, per term i and pose:
    |gpu_i - ref_i| <= 1e-4 * mag_i          (mag_i = sum |pair contribution|)
and, for terms without heavy cancellation (mag_i <= 10 |ref_i|),
    |gpu_i - ref_i| <= 1e-4 * |ref_i|.
Decisions (Metropolis, exchange) and the random stream: bit-exact.
"""
import numpy as np
import pytest

from geaxdock_gpu_amd import abi, synth

pytestmark = pytest.mark.gpu

TERMS = ["vdw", "ele", "pmf", "psp", "hdb", "hpc", "kde", "lhm", "dst", "total"]


@pytest.fixture(scope="module")
def eng_mod():
    import geaxdock_gpu_amd as g
    return g


def random_poses(rng, n, t_sigma=1.5):
    mm = np.zeros((n, 6), np.float32)
    mm[:, :3] = rng.normal(0.0, t_sigma, size=(n, 3))
    mm[:, 3:] = rng.uniform(-np.pi, np.pi, size=(n, 3))
    mm[0] = 0.0
    return mm


def check_energy_parity(oracle, eng, cid, cx, mm, idx_lig, idx_prt, report):
    d = cx.desc()
    ref = oracle.refmap_f64(d, cx)
    out = eng.calc_energy(cid, idx_lig, idx_prt, mm)
    worst_rel = np.zeros(10)
    worst_cond = np.zeros(10)
    for i in range(len(mm)):
        o = oracle.energy_f64(d, ref, int(idx_prt[i]), int(idx_lig[i]), mm[i])
        g = out[i]
        for t in range(10):
            r, m = o.e[t], max(o.mag[t], 1e-30)
            err = abs(float(g["e"][t]) - r)
            cond = err / m
            worst_cond[t] = max(worst_cond[t], cond)
            assert cond <= 1e-4, (TERMS[t], i, float(g["e"][t]), r, m)
            if m <= 10.0 * abs(r) and abs(r) > 0:
                rel = err / abs(r)
                worst_rel[t] = max(worst_rel[t], rel)
                assert rel <= 1e-4, (TERMS[t], i, float(g["e"][t]), r)
        assert abs(float(g["rmsd"]) - o.rmsd) <= 1e-4 * max(1.0, o.rmsd), (i, g["rmsd"], o.rmsd)
        if o.min_margin > 1e-3:
            assert abs(float(g["cms"]) - o.cms) <= 1e-5, (i, g["cms"], o.cms)
    report.update({TERMS[t]: (worst_rel[t], worst_cond[t]) for t in range(10)})
    return out


@pytest.mark.parametrize("lpr,apl,split", [(8, 4, 1), (8, 4, 2), (16, 2, 1), (16, 3, 1),
                                           (16, 3, 2), (32, 1, 2), (64, 1, 1), (16, 2, 4),
                                           (8, 4, 8)])
def test_energy_terms_vs_fp64(oracle, eng_mod, lpr, apl, split):
    cx = synth.single_complex(seed=3, identity_norm=True, void_radius=7.0)
    rng = np.random.default_rng(11)
    mm = random_poses(rng, 48)
    with eng_mod.Engine(seed=1, lanes_per_replica=lpr, atoms_per_lane=apl,
                        receptor_split=split) as eng:
        cid = eng.add_complex(cx)
        rep = {}
        check_energy_parity(oracle, eng, cid, cx, mm, np.zeros(48, np.int32),
                            np.zeros(48, np.int32), rep)
    print("\nlpr", lpr, "apl", apl, "split", split, "worst (rel, cond):",
          {k: (f"{v[0]:.2e}", f"{v[1]:.2e}") for k, v in rep.items()})


def test_energy_terms_weighted_multiconformer(oracle, eng_mod):
    """Normalization a*e+b and weights of the reference parameter file, several
    ligand and protein conformers, a 48-atom ligand."""
    cx = synth.make_complex(5, lna=48, pnp=1500, pnk=150, pos=6, mcs_atoms=30, n_lig=3, n_prt=2,
                            n_tmp=2, void_radius=8.0)
    rng = np.random.default_rng(5)
    n = 36
    mm = random_poses(rng, n, t_sigma=2.0)
    il = rng.integers(0, 3, size=n).astype(np.int32)
    ip = rng.integers(0, 2, size=n).astype(np.int32)
    il[0] = ip[0] = 0
    with eng_mod.Engine(seed=2) as eng:
        cid = eng.add_complex(cx)
        rep = {}
        check_energy_parity(oracle, eng, cid, cx, mm, il, ip, rep)
    print("\nworst (rel, cond):", {k: (f"{v[0]:.2e}", f"{v[1]:.2e}") for k, v in rep.items()})


def test_energy_terms_extreme_pmf_distances(oracle, eng_mod):
    """PMF distances far outside the parameter file's range (pmf0 = 45 for
    receptor type 0, -20 for types 1-7): the PMF exponent (6 - d/2)(d - pmf0)
    reaches 120-136 nats, so exp2 would overflow without its clamp. Those
    chunks run the pair loop with the clamp; the rest run it without (both
    variants against the FP64 oracle). pmf0 = 45 also widens the contact
    radius of type 0 to 45 A, which the contact map and MCC follow (one type
    only, so no ligand atom passes the engine's 256 reference contacts)."""
    cx = synth.single_complex(seed=3, identity_norm=True, void_radius=7.0)
    ep = cx.enepara_struct
    for lt in range(abi.MAXTP2):
        ep.pmf0[lt][0] = 45.0
        for pt in range(1, 8):
            ep.pmf0[lt][pt] = -20.0
    cx.enepara = np.frombuffer(bytes(ep), dtype=np.uint8).copy()
    rng = np.random.default_rng(13)
    mm = random_poses(rng, 32)
    with eng_mod.Engine(seed=1) as eng:
        cid = eng.add_complex(cx)
        rep = {}
        out = check_energy_parity(oracle, eng, cid, cx, mm, np.zeros(32, np.int32),
                                  np.zeros(32, np.int32), rep)
    assert np.isfinite(out["e"]).all()
    print("\nworst (rel, cond):", {k: (f"{v[0]:.2e}", f"{v[1]:.2e}") for k, v in rep.items()})


def test_start_pose_mcc_is_one(oracle, eng_mod):
    """Replica 0 in its start pose reproduces the reference contact map exactly:
    cms = 1 whenever the map has contacts (src/kernel_cuda_l2_calcmcc.cu)."""
    cx = synth.single_complex(seed=4, void_radius=5.5)
    d = cx.desc()
    ref = oracle.refmap_f64(d, cx)
    assert ref.sum() > 0
    with eng_mod.Engine(seed=3) as eng:
        cid = eng.add_complex(cx)
        out = eng.calc_energy(cid, [0], [0], np.zeros((1, 6), np.float32))
    assert float(out[0]["cms"]) == pytest.approx(1.0, abs=1e-6)
    assert float(out[0]["rmsd"]) == 0.0


def test_mcs_templates_on_the_start_pose(oracle, eng_mod):
    """Every MCS template atom placed exactly on the start pose (its FP32
    coordinates): the MCS term is log(0) = -inf, as in the reference
    (src/kernel_cuda_l2_calcenergy.cu:341), and so is the weighted total, on the
    device as on the FP32 CPU path, which places the pose with the same FP32
    additions. (The FP64 oracle places it exactly and sees the FP32 rounding
    of the template, a finite log.) A replica holding -inf accepts no move
    (exp of -inf is 0), so an MC run from the start poses keeps every replica
    there."""
    cx = synth.single_complex(seed=4)
    start = np.stack([cx.lig_x[:cx.lna], cx.lig_y[:cx.lna], cx.lig_z[:cx.lna]], 1) + \
        np.asarray(cx.pocket_center, np.float32)
    for arr, i in ((cx.mcs_x, 0), (cx.mcs_y, 1), (cx.mcs_z, 2)):
        t = arr.reshape(cx.pos, cx.n_mcs)
        valid = t != abi.MCS_INVALID_COORD
        for m in range(cx.pos):
            for a in np.nonzero(valid[m, 1:])[0]:
                t[m, a + 1] = start[a, i]
    d = cx.desc()
    want = oracle.energy_f32(d, oracle.refmap_f32(d, cx), 0, 0, np.zeros(6, np.float32))
    assert want.e[7] == -np.inf and want.e[9] == -np.inf
    with eng_mod.Engine(seed=3, record=False) as eng:
        cid = eng.add_complex(cx)
        out = eng.calc_energy(cid, [0], [0], np.zeros((1, 6), np.float32))
        assert out[0]["e"][7] == -np.inf and out[0]["e"][9] == -np.inf
        eng.run(40, 20)
        st = eng.state(cid)
    assert np.all(st["acs_mc"] == 0)
    assert np.all(st["movematrix"] == 0) and np.all(st["energy"]["e"][:, 9] == -np.inf)


def test_philox_matches_oracle(oracle, eng_mod):
    rng = np.random.default_rng(0)
    ctr = rng.integers(0, 2**32, size=(257, 4), dtype=np.uint64).astype(np.uint32)
    key = rng.integers(0, 2**32, size=(257, 2), dtype=np.uint64).astype(np.uint32)
    ctr[0], key[0] = 0, 0
    ctr[1], key[1] = 0xFFFFFFFF, 0xFFFFFFFF
    with eng_mod.Engine() as eng:
        got = eng.philox(ctr, key)
    for i in range(len(ctr)):
        assert np.array_equal(got[i], oracle.philox(ctr[i], key[i])), i
    assert [hex(v) for v in got[0]] == ["0x6627e8d5", "0xe169c58d", "0xbc57ac4c", "0x9b00dbd8"]


def test_accept_bit_exact(oracle, eng_mod):
    rng = np.random.default_rng(1)
    n = 200000
    e_old = rng.normal(0, 3, n).astype(np.float32)
    e_new = (e_old + rng.normal(0, 1, n)).astype(np.float32)
    mb = (-1.0 / rng.uniform(0.2, 6.0, n)).astype(np.float32)
    u = ((rng.integers(0, 2**24, n) + 1) * 2.0**-24).astype(np.float32)
    # edge cases: dE = 0 with U = 1 (strict U < exp(0) rejects) and U < 1,
    # NaN / inf energies, huge dE
    e_new[:9] = [e_old[0], np.nan, np.inf, -np.inf, 1e30, -1e30, e_old[6], e_old[7] + 1e-7,
                 e_old[8]]
    u[:9] = [1.0, 0.5, 0.5, 0.5, 0.5, 0.5, 2.0**-24, 1.0, 1.0 - 2.0**-24]
    with eng_mod.Engine() as eng:
        got = eng.accept(e_new, e_old, mb, u)
    ref = np.array([oracle.lib().gdo_accept(float(a), float(b), float(c), float(d))
                    for a, b, c, d in zip(e_new, e_old, mb, u)], np.int32)
    assert np.array_equal(got, ref)
    assert got[0] == 0 and got[8] == 1  # dE = 0: accept unless U = 1
    assert got[1] == 0 and got[2] == 0  # NaN and +inf never accepted
    assert got[3] == 1 and got[5] == 1 and got[4] == 0


def test_exchange_bit_exact(oracle, eng_mod):
    import ctypes as C
    rng = np.random.default_rng(2)
    for n_tmp in (2, 7, 8, 16, 40):
        n_groups = 500
        et = rng.normal(0, 2, size=(n_groups, n_tmp)).astype(np.float32)
        it = np.array([rng.permutation(n_tmp) for _ in range(n_groups)], np.int32)
        ac = rng.integers(0, 5, size=(n_groups, n_tmp)).astype(np.int32)
        _, mb = synth.temperature_ladder(n_tmp, 0.5, 5.0)
        u = rng.uniform(0, 1, size=(n_groups, n_tmp // 2)).astype(np.float32)
        for mode_t in (0, 1):
            with eng_mod.Engine() as eng:
                git, gac = eng.exchange(n_tmp, mode_t, et, mb, u, it, ac)
            for g in range(n_groups):
                rit, rac = it[g].copy(), ac[g].copy()
                oracle.lib().gdo_exchange_group(
                    n_tmp, mode_t, et[g].ctypes.data_as(C.POINTER(C.c_float)),
                    mb.ctypes.data_as(C.POINTER(C.c_float)),
                    u[g].ctypes.data_as(C.POINTER(C.c_float)),
                    rit.ctypes.data_as(C.POINTER(C.c_int32)),
                    rac.ctypes.data_as(C.POINTER(C.c_int32)))
                assert np.array_equal(git[g], rit) and np.array_equal(gac[g], rac), (n_tmp, mode_t, g)
                assert sorted(git[g].tolist()) == list(range(n_tmp))


def _run_gpu(eng_mod, cx, seed, n_steps, nc, **kw):
    recs = []
    with eng_mod.Engine(seed=seed, **kw) as eng:
        cid = eng.add_complex(cx)
        st = eng.run(n_steps, nc, sink=lambda c, r: recs.append(r))
        state = eng.state(cid)
    recs = np.concatenate(recs) if recs else np.zeros(0, abi.RECORD_DTYPE)
    return recs, state, st


@pytest.mark.parametrize("split", [1, 2, 4])
def test_mc_run_tracks_cpu_path(oracle, eng_mod, split):
    """Same seed, same random stream: the GPU engine and the FP32 CPU path make
    the same accept / exchange decisions, step for step. A replica may diverge
    only where the CPU path's own margin log shows a decision of it or its
    ladder within FP32 rounding of its threshold (oracle/decisions.py); every
    other replica's counters and temperature index must be equal. One, two
    and four receptor-split groups."""
    from oracle import decisions
    cx = synth.single_complex(seed=0)
    n_steps, nc = 400, 20
    recs, state, st = _run_gpu(eng_mod, cx, 7, n_steps, nc, receptor_split=split)
    d = cx.desc()
    cpu = oracle.McState(cx, rec_cap=n_steps + 1)
    cpu.enable_log(0, n_steps, n_steps // nc)
    oracle.mc_run(d, cpu, seed=7, init=True, step0=0, n_steps=0, nc=nc)
    oracle.mc_run(d, cpu, seed=7, init=False, step0=0, n_steps=n_steps, nc=nc)
    out = decisions.compare(cx, recs, cpu, nc, gpu_state=state)
    print("\nidentical decision sequences:", out["identical"], "of", cx.n_rep, "explained:",
          out["explained"], "dE", out["d_e"], "acs_mc", state["acs_mc"], "acs_x",
          state["acs_x"])


@pytest.mark.parametrize("case", ["plain", "plain_split4", "da"])
def test_uniform_moves_give_the_cpu_poses_bit_for_bit(oracle, eng_mod, case):
    """With the reference's uniform moves (move_scale * (2U - 1)) every pose
    the engine accepts, up to a replica's first explained divergence, has the
    bits of the CPU path's: the same Philox uniforms, and one product and one
    sum per coordinate, rounded one at a time on both sides."""
    from oracle import decisions
    from geaxdock_gpu_amd import surrogate as sg
    cx = synth.make_complex(seed=5, lna=32, pnp=1200, n_lig=2, n_tmp=8,
                            move_kind=abi.MOVE_UNIFORM)
    n_steps, nc, k = 200, 20, 4
    da = case == "da"
    recs = []
    kw = dict(receptor_split=4) if case == "plain_split4" else {}
    with eng_mod.Engine(seed=11, **kw) as eng:
        cid = eng.add_complex(cx)
        if da:
            eng.sur_build_grids(0.8, 12.0)
            eng.sur_set_params(sg.pack_params(sg.init_params(3)))
            eng.run_da(n_steps, nc, k, 0.0, sink=lambda c, r: recs.append(r))
        else:
            eng.run(n_steps, nc, sink=lambda c, r: recs.append(r))
        state = eng.state(cid)
    recs = np.concatenate(recs)
    d = cx.desc()
    cpu = oracle.McState(cx, rec_cap=n_steps + 1)
    cpu.enable_log(0, n_steps, n_steps // nc)
    oracle.mc_run(d, cpu, seed=11, init=True, step0=0, n_steps=0, nc=nc)
    if da:
        oracle.mc_run_da(d, cpu, seed=11, step0=0, n_steps=n_steps, nc=nc, k_sub=k)
    else:
        oracle.mc_run(d, cpu, seed=11, init=False, step0=0, n_steps=n_steps, nc=nc)
    out = decisions.compare(cx, recs, cpu, nc, k_sub=k if da else 1, gpu_state=state, da=da)
    print(f"\n{case}: {out['identical']}/{cx.n_rep} identical; {out['poses']} accepted poses, "
          f"{out['poses_differing']} differing")
    assert out["poses"] > 10 * cx.n_rep and out["poses_differing"] == 0


def test_checkpoint_resume_bit_exact(eng_mod):
    cx = synth.single_complex(seed=1)
    with eng_mod.Engine(seed=5, record=False) as eng:
        cid = eng.add_complex(cx)
        eng.run(60, 20)
        blob = eng.checkpoint()
        eng.run(80, 20)
        a = eng.state(cid)
        eng.restore(blob)
        eng.run(80, 20)
        b = eng.state(cid)
    for k in a:
        assert np.array_equal(a[k], b[k]), k


def test_run_deterministic_and_window_independent(eng_mod):
    """A run is a function of the seed alone: record windows of 50 or 2000
    steps give bit-identical states."""
    cx = synth.single_complex(seed=2)
    _, s1, _ = _run_gpu(eng_mod, cx, 9, 300, 20, steps_per_dump=50)
    _, s2, _ = _run_gpu(eng_mod, cx, 9, 300, 20, steps_per_dump=2000)
    for k in s1:
        assert np.array_equal(s1[k], s2[k]), k


def test_record_drains_independent_of_windows_and_resume(eng_mod):
    """Records reach the sink the same whatever the window: 50-step windows
    (drained from alternating rings while the next window runs) and one
    300-step window deliver the same records, in the same order per replica;
    records after a checkpoint equal those after restoring it."""
    cx = synth.single_complex(seed=6)

    def collect(dump, steps, eng=None):
        out = []
        own = eng is None
        if own:
            eng = eng_mod.Engine(seed=13, steps_per_dump=dump)
            eng.add_complex(cx)
        eng.run(steps, 20, sink=lambda c, r: out.append(r))
        recs = np.concatenate(out)
        order = np.lexsort((recs["step"], recs["replica"][:, 0]))
        if own:
            eng.close()
        return recs[order]

    a = collect(50, 300)
    b = collect(2000, 300)
    assert len(a) == len(b) and len(a) > 300
    assert a.tobytes() == b.tobytes()
    with eng_mod.Engine(seed=13, steps_per_dump=40) as eng:
        eng.add_complex(cx)
        collect(40, 100, eng)
        blob = eng.checkpoint()
        r1 = collect(40, 120, eng)
        eng.restore(blob)
        r2 = collect(40, 120, eng)
    assert len(r1) > 0 and r1.tobytes() == r2.tobytes()


def _batch_states(eng_mod, cxs, seed, n_steps, nc, **kw):
    with eng_mod.Engine(seed=seed, record=False, **kw) as eng:
        for cx in cxs:
            eng.add_complex(cx)
        eng.run(n_steps, nc)
        return [eng.state(c) for c in range(len(cxs))]


def test_launch_modes_bit_identical(eng_mod):
    """Graph-captured, eager, persistent and non-persistent launches of an
    uneven batch give bit-identical states (graph replays included)."""
    cxs = synth.uneven_batch(3, seed=5, lna_range=(12, 40), pnp_range=(800, 1600))
    ref = None
    for launch in (0, abi.LAUNCH_NO_GRAPH, abi.LAUNCH_PERSISTENT,
                   abi.LAUNCH_PERSISTENT | abi.LAUNCH_NO_GRAPH):
        st = _batch_states(eng_mod, cxs, 3, 120, 20, launch=launch, steps_per_dump=40)
        if ref is None:
            ref = st
            continue
        for a, b in zip(ref, st):
            for k in a:
                assert np.array_equal(a[k], b[k]), (launch, k)


@pytest.mark.parametrize("split,ladders", [(1, 0), (2, 1), (1, 3), (0, 0), (4, 2), (8, 1)])
def test_batch_tracks_cpu_path(oracle, eng_mod, split, ladders):
    """An uneven batch in one engine: every complex follows the FP32 CPU path
    of that complex (its own Philox streams, keyed by complex id), decision
    for decision up to explained rounding splits (oracle/decisions.py)."""
    from oracle import decisions
    cxs = synth.uneven_batch(3, seed=6, lna_range=(12, 40), pnp_range=(800, 1600), n_lig=2,
                             n_prt=2, n_tmp=4)
    n_steps, nc = 200, 20
    recs = {c: [] for c in range(len(cxs))}
    with eng_mod.Engine(seed=4, receptor_split=split, ladders_per_pack=ladders) as eng:
        for cx in cxs:
            eng.add_complex(cx)
        eng.run(n_steps, nc, sink=lambda c, r: recs[c].append(r))
        states = [eng.state(c) for c in range(len(cxs))]
    total = same = 0
    for c, cx in enumerate(cxs):
        g_all = np.concatenate(recs[c])
        d = cx.desc()
        cpu = oracle.McState(cx, rec_cap=n_steps + 1)
        cpu.enable_log(0, n_steps, n_steps // nc)
        oracle.mc_run(d, cpu, seed=4, cid=c, init=True, step0=0, n_steps=0, nc=nc)
        oracle.mc_run(d, cpu, seed=4, cid=c, init=False, step0=0, n_steps=n_steps, nc=nc)
        out = decisions.compare(cx, g_all, cpu, nc, gpu_state=states[c])
        total += cx.n_rep
        same += out["identical"]
        if out["explained"]:
            print(f"\ncomplex {c}: explained {out['explained']}")
    print(f"\nsplit {split} ladders {ladders}: {same}/{total} replicas with identical decisions")


@pytest.mark.parametrize("split,ladders", [(0, 0), (0, 4), (1, 0)])
def test_odd_ladder_shapes_track_cpu_path(oracle, eng_mod, split, ladders):
    """Ladders of 1 to 5 waves (3, 6, 12, 20 and 40 temperatures; 8 or 16 lanes
    per replica) and odd ligand-conformer counts, so the automatic split makes
    packs of 1, 2 and 4 ladders at splits 1 to 8, with 3- and 5-wave groups:
    every complex follows the FP32 CPU path decision for decision (up to
    explained rounding splits), records equal gd_calc_energy of their poses bit
    for bit, and graph-captured persistent and eager launches agree."""
    from oracle import decisions
    shapes = [(3, 20, 5), (6, 40, 3), (12, 36, 3), (20, 14, 2), (40, 12, 1)]
    cxs = [synth.make_complex(50 + i, lna=lna, pnp=700 + 150 * i, n_lig=n_lig, n_prt=1,
                              n_tmp=n_tmp)
           for i, (n_tmp, lna, n_lig) in enumerate(shapes)]
    n_steps, nc = 120, 20
    states = {}
    for launch in (abi.LAUNCH_PERSISTENT, abi.LAUNCH_NO_GRAPH):
        recs = {c: [] for c in range(len(cxs))}
        with eng_mod.Engine(seed=9, receptor_split=split, ladders_per_pack=ladders,
                            launch=launch, steps_per_dump=60) as eng:
            for cx in cxs:
                eng.add_complex(cx)
            eng.run(n_steps, nc, sink=lambda c, r: recs[c].append(r))
            states[launch] = [eng.state(c) for c in range(len(cxs))]
            if launch != abi.LAUNCH_PERSISTENT:
                continue
            for c, cx in enumerate(cxs):
                r = np.concatenate(recs[c])
                ex = eng.calc_energy(c, r["replica"][:, 3], r["replica"][:, 1], r["movematrix"])
                assert np.array_equal(ex["e"].view(np.uint32), r["energy"]["e"].view(np.uint32))
                d = cx.desc()
                cpu = oracle.McState(cx, rec_cap=n_steps + 1)
                cpu.enable_log(0, n_steps, n_steps // nc)
                oracle.mc_run(d, cpu, seed=9, cid=c, init=True, step0=0, n_steps=0, nc=nc)
                oracle.mc_run(d, cpu, seed=9, cid=c, init=False, step0=0, n_steps=n_steps, nc=nc)
                out = decisions.compare(cx, r, cpu, nc, gpu_state=states[launch][c])
                print(f"\nsplit {split} ladders {ladders} complex {c} (n_tmp {cx.n_tmp}, "
                      f"lna {cx.lna}): "
                      f"{out['identical']}/{cx.n_rep} identical, explained {out['explained']}")
    for a, b in zip(states[abi.LAUNCH_PERSISTENT], states[abi.LAUNCH_NO_GRAPH]):
        for k in a:
            assert np.array_equal(a[k], b[k]), k


def test_statistics_match_cpu_path_independent_seeds(oracle, eng_mod):
    """Independent chains (different seeds) of the GPU engine and the FP32 CPU
    path on one complex, exchanges off: at every temperature the per-replica
    acceptance ratios and minimum energies come from the same distributions
    (two-sample Kolmogorov-Smirnov, p > 1e-4 for each of the 8 comparisons)."""
    from scipy.stats import ks_2samp
    cx = synth.make_complex(seed=7, lna=24, pnp=900, n_lig=32, n_prt=1, n_tmp=4,
                            t_floor=0.5, t_ceiling=5.0)
    n_steps = 400
    with eng_mod.Engine(seed=101, record=False) as eng:
        cid = eng.add_complex(cx)
        eng.run(n_steps, 0)
        st = eng.state(cid)
        g_best = eng.best(cid)[0].astype(np.float64)
    g_acc = st["acs_mc"] / n_steps
    d = cx.desc()
    cpu = oracle.McState(cx, rec_cap=n_steps + 1)
    oracle.mc_run(d, cpu, seed=202, init=True, step0=0, n_steps=0, nc=0)
    oracle.mc_run(d, cpu, seed=202, init=False, step0=0, n_steps=n_steps, nc=0)
    c_acc = cpu.acs_mc / n_steps
    c_best = np.array([cpu.records(r)["energy"]["e"][:, 9].min() for r in range(cx.n_rep)],
                      np.float64)
    temp = (np.arange(cx.n_rep) // cx.n_lig) % cx.n_tmp
    for t in range(cx.n_tmp):
        sel = temp == t
        p_acc = ks_2samp(g_acc[sel], c_acc[sel]).pvalue
        p_min = ks_2samp(g_best[sel], c_best[sel]).pvalue
        print(f"\nT{t}: acceptance GPU {g_acc[sel].mean():.4f} CPU {c_acc[sel].mean():.4f} "
              f"p={p_acc:.3f}; min energy GPU {g_best[sel].mean():.4f} CPU "
              f"{c_best[sel].mean():.4f} p={p_min:.3f}")
        assert p_acc > 1e-4 and p_min > 1e-4, (t, p_acc, p_min)
