"""Delayed-acceptance MC (gd_run_da) on the GPU.

* Surrogate scale 0 reduces the surrogate to the exact cheap terms; the chain
  then has a CPU restatement (oracle gdo_mc_run_da) and the decision
  sequences must agree step for step, as for the plain path.
* With a real (random or trained) MLP every recorded state and the final
  state carry the exact energy of their pose (FP32, 1e-4 of the pair
  magnitude), a run is a function of the seed alone (record windows of any
  length give bit-identical states), and an exact rejection never moves a
  replica.
* The composite test keeps the exact Boltzmann distribution: with exchanges
  off, the per-temperature mean energy of a DA chain matches that of the
  plain chain within sampling error.
"""
import numpy as np
import pytest

from geaxdock_gpu_amd import abi, synth
from geaxdock_gpu_amd import surrogate as sg

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gd():
    import geaxdock_gpu_amd as g
    return g


def _engine(gd, cx, seed, params_seed=3, tab=False, **kw):
    eng = gd.Engine(seed=seed, **kw)
    cid = eng.add_complex(cx)
    eng.sur_build_grids(0.8, 12.0)
    eng.sur_set_params(sg.pack_params(sg.init_params(params_seed)))
    if tab:
        eng.sur_tabulate()
    return eng, cid


# A complex with large KDE and MCS sets: the groups of a receptor split share
# the surrogate's KDE and MCS sums (under spread packing, split 8, and under a
# fixed split of 4).
def _large_lists():
    return synth.make_complex(seed=31, lna=24, pnp=900, pnk=320, pos=36, mcs_atoms=20, n_lig=2,
                              n_prt=1, n_tmp=6)


@pytest.mark.parametrize("case", ["single", "shared_spread", "shared_split4"])
def test_da_scale0_tracks_cpu_path(gd, case):
    """Decision for decision against the CPU DA chain, up to explained rounding
    splits of a sub-step, composite or exchange test (oracle/decisions.py)."""
    from oracle import decisions, gdo
    cx = synth.single_complex(seed=0) if case == "single" else _large_lists()
    kw = dict(receptor_split=4) if case == "shared_split4" else {}
    n_steps, nc, k = 400, 20, 4
    recs = []
    eng, cid = _engine(gd, cx, 11, **kw)
    st = eng.run_da(n_steps, nc, k, 0.0, sink=lambda c, r: recs.append(r))
    state = eng.state(cid)
    eng.close()
    recs = np.concatenate(recs)
    d = cx.desc()
    cpu = gdo.McState(cx, rec_cap=n_steps + 1)
    cpu.enable_log(0, n_steps, n_steps // nc)
    gdo.mc_run(d, cpu, seed=11, init=True, step0=0, n_steps=0, nc=nc)
    gdo.mc_run_da(d, cpu, seed=11, step0=0, n_steps=n_steps, nc=nc, k_sub=k)
    out = decisions.compare(cx, recs, cpu, nc, k_sub=k, gpu_state=state, da=True)
    print("\nDA(scale 0) replicas with identical decisions:", out["identical"], "of", cx.n_rep,
          "explained", out["explained"], "acs", state["acs_mc"], "x", state["acs_x"])
    assert st["accepted"] == int(state["acs_mc"].sum())
    # composite moves are recorded at the last sub-step of a group
    assert np.all((recs["step"][recs["step"] > 0] + 1) % k == 0)


@pytest.mark.parametrize("tab", [False, True])
def test_da_states_carry_exact_energy(gd, tab):
    cx = synth.single_complex(seed=1)
    recs = []
    eng, cid = _engine(gd, cx, 5, tab=tab)
    eng.run_da(240, 20, 4, 3.0, sink=lambda c, r: recs.append(r))
    state = eng.state(cid)
    recs = np.concatenate(recs)
    ex = eng.calc_energy(cid, recs["replica"][:, 3], recs["replica"][:, 1], recs["movematrix"])
    np.testing.assert_allclose(recs["energy"]["e"], ex["e"], rtol=1e-5, atol=1e-5)
    ex = eng.calc_energy(cid, state["replica"][:, 3], state["replica"][:, 1],
                         state["movematrix"])
    np.testing.assert_allclose(state["energy"]["e"], ex["e"], rtol=1e-5, atol=1e-5)
    # every replica's last record is its current state
    for r in range(cx.n_rep):
        last = recs[recs["replica"][:, 0] == r][-1]
        np.testing.assert_array_equal(last["movematrix"], state["movematrix"][r])
    eng.close()


@pytest.mark.parametrize("tab", [False, True])
def test_da_deterministic_and_window_independent(gd, tab):
    cx = synth.single_complex(seed=2)
    out = []
    for dump in (40, 2000):
        eng, cid = _engine(gd, cx, 9, tab=tab, steps_per_dump=dump)
        eng.run_da(200, 20, 4, 2.0)
        out.append(eng.state(cid))
        eng.close()
    for key in out[0]:
        assert np.array_equal(out[0][key], out[1][key]), key


def test_da_rejects_bad_arguments(gd):
    cx = synth.single_complex(seed=2)
    with gd.Engine(seed=1) as eng:
        eng.add_complex(cx)
        with pytest.raises(gd.GeauxDockError):
            eng.run_da(40, 20, 4, 1.0)  # no grids yet
        eng.sur_build_grids(0.8, 12.0)
        with pytest.raises(gd.GeauxDockError):
            eng.run_da(40, 20, 4, 1.0)  # no parameters yet
        eng.sur_set_params(sg.pack_params(sg.init_params(0)))
        with pytest.raises(gd.GeauxDockError):
            eng.run_da(42, 20, 4, 1.0)  # not a multiple of k
        with pytest.raises(gd.GeauxDockError):
            eng.run_da(40, 10, 4, 1.0)
        eng.run_da(40, 20, 4, 1.0)


def test_da_keeps_boltzmann_mean_energy(gd):
    """Plain and DA chains (no exchange) at fixed temperatures: after burn-in,
    per-temperature mean total energies agree within 5 standard errors of
    the mean paired difference over replicas (independent chains)."""
    cx = synth.make_complex(seed=4, lna=24, pnp=900, n_lig=16, n_prt=1, n_tmp=4,
                            t_floor=1.0, t_ceiling=4.0)
    k = 4
    means = []
    for da in (False, True):
        eng, cid = _engine(gd, cx, 21, tab=True, record=False)
        run = (lambda n: eng.run_da(n, 0, k, 1.0)) if da else (lambda n: eng.run(n, 0))
        run(3000)
        acc = []
        for _ in range(30):
            run(100)
            s = eng.state(cid)
            acc.append(s["energy"]["e"][:, 9].astype(np.float64))
        eng.close()
        e = np.stack(acc)  # [sample, replica]
        it = s["replica"][:, 2]
        means.append([e[:, it == t].mean(0) for t in range(cx.n_tmp)])
    for t in range(cx.n_tmp):
        # per-replica time means, paired by replica (same ligand conformer)
        dlt = means[0][t] - means[1][t]
        se = dlt.std(ddof=1) / np.sqrt(len(dlt))
        print(f"T{t}: plain {means[0][t].mean():.4f} DA {means[1][t].mean():.4f} se {se:.4f}")
        assert abs(dlt.mean()) <= 5 * se + 1e-3, (t, dlt.mean(), se)


@pytest.mark.parametrize("tab", [False, True])
def test_surrogate_total_matches_oracle(gd, tab):
    """S = scale * sum_atoms y / lna + weighted exact KDE, MCS and pocket
    distance, as the delayed-acceptance kernel computes it, against the FP64
    network (oracle/surrogate.py) on the FP64 feature grid. Direct mode
    interpolates the features and evaluates the network per atom; tabulated
    mode evaluates the network at the grid nodes and interpolates its
    output. BF16 operands bound the agreement: |dS| <= 2e-2 * scale *
    sum|y| / lna + 1e-4 (1 + |cheap|)."""
    from oracle import gdo
    from oracle import surrogate as so
    cx = synth.single_complex(seed=2, void_radius=7.0)
    spacing, half, scale = 0.8, 9.6, 7.0
    p = so.init_params(3)
    with gd.Engine(seed=1) as eng:
        cid = eng.add_complex(cx)
        eng.sur_build_grids(spacing, half)
        eng.sur_set_params(sg.pack_params(p))
        if tab:
            eng.sur_tabulate()
        rng = np.random.default_rng(4)
        n_pose = 16
        mm = np.zeros((n_pose, 6), np.float32)
        mm[:, :3] = rng.normal(0, 1.5, (n_pose, 3))
        mm[:, 3:] = rng.uniform(-np.pi, np.pi, (n_pose, 3))
        got = eng.sur_surrogate(cid, np.zeros(n_pose), np.zeros(n_pose), mm, scale)
    n = int(np.ceil(2 * half / spacing)) + 1
    origin = np.float64(cx.pocket_center) - half
    prt = np.stack([cx.prt_x, cx.prt_y, cx.prt_z], 1)[:cx.pnp]
    grid = so.grid_features(origin, spacing, n, prt, cx.prt_t)
    tabs = so.tabulate(p, grid, np.unique(cx.lig_t)) if tab else None
    d = cx.desc()
    ref = gdo.refmap_f64(d, cx)
    w = np.asarray(cx.enepara_struct.w, np.float64)
    worst = 0.0
    for i in range(n_pose):
        xyz = gdo.place_f64(d, cx, 0, mm[i]).T
        if tab:
            y = np.array([so.interp(tabs[int(t)][..., None], origin, spacing, xyz[k:k + 1])[0, 0]
                          for k, t in enumerate(cx.lig_t)])
        else:
            y = so.rows(p, so.features(cx.lig_t, so.interp(grid, origin, spacing, xyz)))
        o = gdo.energy_f64(d, ref, 0, 0, mm[i])
        cheap = sum(w[j] * o.e[j] for j in (6, 7, 8))
        want = scale * y.sum() / cx.lna + cheap
        tol = 2e-2 * scale * np.abs(y).sum() / cx.lna + 1e-4 * (1 + abs(cheap))
        worst = max(worst, abs(got[i] - want) / tol)
        assert abs(got[i] - want) <= tol, (i, got[i], want, tol)
    print(f"\ntab={tab}: worst |dS|/tol = {worst:.3f}")


@pytest.mark.parametrize("tab", [False, True])
def test_da_equals_plain_when_surrogate_is_exact(gd, tab):
    """Receptor-pair weights zero: the exact energy is the surrogate's cheap
    part (scale 0), every composite move is accepted, and the DA engine passes
    through the plain engine's states every k steps, exchanges included
    (rounding between the two code paths may split at most one replica)."""
    w = [0.0] * 6 + [0.7, 0.4, 1.0, 0.0]
    cx = synth.make_complex(seed=12, lna=12, pnp=300, n_lig=4, n_prt=2, n_tmp=4, box=20.0,
                            void_radius=6.0, weights=w, t_floor=0.5, t_ceiling=5.0)
    states = []
    for da in (False, True):
        eng, cid = _engine(gd, cx, 9, tab=tab, record=False)
        if da:
            eng.run_da(400, 20, 4, 0.0)
        else:
            eng.run(400, 20)
        states.append(eng.state(cid))
        eng.close()
    p, q = states
    # the unweighted pair terms are compared loosely: their last bits depend
    # on which far-field range the wave's other poses put each point in
    same = np.all(p["movematrix"] == q["movematrix"], axis=1) & \
        np.all(p["energy"]["e"][:, 6:] == q["energy"]["e"][:, 6:], axis=1)
    np.testing.assert_allclose(p["energy"]["e"][same, :6], q["energy"]["e"][same, :6],
                               rtol=1e-5, atol=1e-6)
    same_mm = np.all(p["movematrix"] == q["movematrix"], axis=1)
    print(f"\ntab={tab}: {same.sum()} of {len(same)} replicas identical "
          f"({same_mm.sum()} same pose); plain acceptance {p['acs_mc'].mean() / 400:.3f}")
    assert same.sum() >= len(same) - 1, same
    assert np.array_equal(p["replica"][:, 2], q["replica"][:, 2]) or same.sum() < len(same)


@pytest.mark.parametrize("tab,shared", [(False, False), (True, False), (False, True),
                                       (True, True)])
def test_surrogate_cheap_terms_equal_exact_terms(gd, tab, shared):
    """With scale 0 the surrogate is its cheap part: the weighted KDE, MCS and
    pocket-distance terms. Its MCS restraint comes in closed form from the
    conformer's template moments (A + n|R obar + t - mbar|^2 - 2<R, M>); it
    must equal the exact per-atom terms of gd_calc_energy to FP32 rounding,
    over many conformers, both receptor models and poses far from and near
    the templates. With `shared`, a complex with large KDE and MCS sets under
    a receptor split of 4: the split's groups add up the sums together."""
    pnk, pos, kw = (320, 36, dict(receptor_split=4)) if shared else (120, 7, {})
    cx = synth.make_complex(seed=21, lna=40, pnp=600, pnk=pnk, pos=pos, mcs_atoms=30, n_lig=5,
                            n_prt=2, n_tmp=2)
    rng = np.random.default_rng(4)
    n = 300
    mm = np.zeros((n, 6), np.float32)
    mm[:, :3] = rng.normal(0, 2.0, (n, 3)) * (rng.random((n, 1)) < 0.7)
    mm[:, 3:] = rng.uniform(-np.pi, np.pi, (n, 3)) * (rng.random((n, 1)) < 0.7)
    il = rng.integers(0, cx.n_lig, n).astype(np.int32)
    ip = rng.integers(0, cx.n_prt, n).astype(np.int32)
    eng, cid = _engine(gd, cx, 3, tab=tab, **kw)
    s = eng.sur_surrogate(cid, il, ip, mm, 0.0).astype(np.float64)
    e = eng.calc_energy(cid, il, ip, mm)["e"].astype(np.float64)
    eng.close()
    w = np.array(cx.enepara_struct.w[:], np.float64)
    want = (e[:, 6:9] * w[6:9]).sum(axis=1)
    err = np.abs(s - want)
    print(f"\ntab={tab} shared={shared}: worst |S - exact cheap terms| {err.max():.2e} (scale {np.abs(want).max():.2f})")
    assert np.all(err <= 2e-5 * np.maximum(1.0, np.abs(want))), err.max()


def test_trained_surrogate_gains_over_no_surrogate(gd):
    """What the learned part buys. Delayed acceptance pays, at its exact
    stage, for the part of a composite move's energy change the screen did
    not see: exp(-beta [(E_y - E_x) - (S_y - S_x)]). Without the network
    (scale 0) the screen sees only the cheap terms, so the residual is the
    whole receptor-pair change; the trained network must explain a good part
    of it on held-out composite moves (k = 4 proposals from logged states),
    while an untrained network only adds noise. (Trained on too few poses
    the network is worse than none: its error on small differences then
    exceeds the differences.)"""
    cx = synth.make_complex(seed=31, lna=32, pnp=1200, n_lig=4, n_prt=2, n_tmp=8)
    k, nc = 4, 20
    eng = gd.Engine(seed=7, record=False)
    cid = eng.add_complex(cx)
    eng.run(200, nc)
    recs, rcid = sg.log_states(eng, 20, 20, nc)
    eng.sur_build_grids(0.6, 12.0)
    ts = sg.build_training_set(eng, recs, rcid, 200000, seed=1)
    _, sigma = sg.load_training_set(eng, ts)
    fit = sg.fit(eng, ts, sigma, n_train=int(0.9 * len(ts["t"])), epochs=20)
    trained = eng.sur_get_params()
    # held-out composite moves: fresh states of a continued run, k steps each
    more, _ = sg.log_states(eng, 4, 20, nc)
    rng = np.random.default_rng(9)
    x = more["movematrix"].astype(np.float32)
    scale = np.array(cx.move_scale, np.float32)
    y = x + (rng.normal(size=(len(x), k, 6)) * scale).sum(axis=1).astype(np.float32)
    il = more["replica"][:, 3].astype(np.int32)
    ip = more["replica"][:, 1].astype(np.int32)
    dE = (eng.calc_energy(cid, il, ip, y)["e"][:, 9] -
          eng.calc_energy(cid, il, ip, x)["e"][:, 9]).astype(np.float64)

    def residual_var(params, s):
        eng.sur_set_params(params)
        eng.sur_tabulate()
        dS = (eng.sur_surrogate(cid, il, ip, y, s) -
              eng.sur_surrogate(cid, il, ip, x, s)).astype(np.float64)
        return float(np.var(dE - dS))

    v_none = residual_var(trained, 0.0)
    v_trained = residual_var(trained, sigma)
    v_random = residual_var(sg.pack_params(sg.init_params(5)), sigma)
    eng.close()
    print(f"\nheld-out RMSE {fit['test_rmse_energy']:.3f} (std {fit['test_energy_std']:.3f}); "
          f"{len(x)} composite moves: residual variance without network {v_none:.4g}, "
          f"trained {v_trained:.4g} ({1 - v_trained / v_none:.0%} explained), "
          f"untrained {v_random:.4g}")
    assert v_trained < 0.6 * v_none
    assert v_random > v_none


def test_screen_where_plain_acceptance_is_low(gd):
    """A cold ladder (kT 0.05-0.5) with large moves (0.8 A, 12 degrees): plain
    Metropolis accepts well under half its proposals. Delayed acceptance from
    one start state with the trained network as the screen accepts more
    composite moves than with the screen reduced to the exact cheap terms
    (scale 0), at the same exact evaluations."""
    cx = synth.make_complex(seed=41, lna=32, pnp=1200, n_lig=4, n_prt=2, n_tmp=8, t_floor=0.05,
                            t_ceiling=0.5, trans_sigma=0.8, rot_sigma_deg=12.0)
    k, nc, n_steps = 4, 20, 800
    eng = gd.Engine(seed=7, record=False)
    cid = eng.add_complex(cx)
    eng.run(200, nc)
    st = eng.run(400, nc)
    plain_acc = st["accepted"] / st["replica_steps"]
    recs, rcid = sg.log_states(eng, 20, 20, nc)
    eng.sur_build_grids(0.6, 12.0)
    ts = sg.build_training_set(eng, recs, rcid, 200000, seed=1)
    _, sigma = sg.load_training_set(eng, ts)
    sg.fit(eng, ts, sigma, n_train=int(0.9 * len(ts["t"])), epochs=20)
    eng.sur_tabulate()
    blob = eng.checkpoint()
    acc = {}
    for tag, scale in (("network", sigma), ("none", 0.0)):
        eng.restore(blob)
        st = eng.run_da(n_steps, nc, k, scale)
        acc[tag] = st["accepted"] / (st["replica_steps"] / k)
    eng.close()
    print(f"\nplain acceptance {plain_acc:.3f}; composite acceptance with network "
          f"{acc['network']:.3f}, without {acc['none']:.3f}")
    assert plain_acc < 0.5
    assert acc["network"] > acc["none"] + 0.01
