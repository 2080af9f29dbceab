"""Every benchmark configuration at its full size on the GPU, through the C ABI.

* Ensemble (3,520 replicas, lna 48, pnp 3,000, 100 MC steps with exchanges):
  10,000 poses (accepted states and random poses) against the FP64 oracle,
  term by term (oracle/parity.py); every recorded energy equal, bit for bit,
  to gd_calc_energy of its pose; every ladder a permutation of temperatures.
* Uneven batch (256 complexes, 30,720 replicas): the same checks per complex
  on sampled poses; graph-captured persistent and eager per-pack launches
  give bit-identical states.
* Surrogate training on 2,000,000 poses drawn from ensemble and batch
  trajectories (20 epochs, batch 4,096, Adam lr 1e-3): held-out RMSE.
* Delayed acceptance over 1,024 complexes (122,880 replicas): recorded and
  final states carry exact energies of their poses (bit-identical to
  gd_calc_energy), runs are independent of the record window, ladders stay
  permutations and the acceptance counters match the records.
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


def _say(msg):
    print(msg, flush=True)


def _ladders_are_permutations(cx, state):
    flat = np.arange(cx.n_rep)
    key = (flat // (cx.n_tmp * cx.n_lig)) * cx.n_lig + flat % cx.n_lig
    it = state["replica"][:, 2]
    for g in np.unique(key):
        assert np.array_equal(np.sort(it[key == g]), np.arange(cx.n_tmp)), g


def _records_are_exact(eng, cid, recs):
    """Recorded energies are those gd_calc_energy gives for the recorded
    poses, bit for bit (same lane layout and receptor split)."""
    ex = eng.calc_energy(cid, recs["replica"][:, 3], recs["replica"][:, 1], recs["movematrix"])
    assert np.array_equal(ex["e"].view(np.uint32), recs["energy"]["e"].view(np.uint32))
    assert np.array_equal(ex["cms"].view(np.uint32), recs["energy"]["cms"].view(np.uint32))


def _fp64(cx, iprt, ilig, mm):
    from oracle import gdo
    d = cx.desc()
    return gdo.energy_f64_batch(d, gdo.refmap_f64(d, cx), iprt, ilig, mm)


def _random_poses(rng, cx, n, t_sigma=1.5):
    mm = np.zeros((n, 6), np.float32)
    mm[:, :3] = rng.normal(0.0, t_sigma, (n, 3))
    mm[:, 3:] = rng.uniform(-np.pi, np.pi, (n, 3))
    return (mm, rng.integers(0, cx.n_prt, n).astype(np.int32),
            rng.integers(0, cx.n_lig, n).astype(np.int32))


def test_ensemble_full_size(gd):
    from oracle import parity
    cx = synth.ensemble(seed=1)
    _say(f"\nensemble: {cx.n_rep} replicas, lna {cx.lna}, pnp {cx.pnp}")
    recs = []
    with gd.Engine(seed=3, steps_per_dump=100) as eng:
        cid = eng.add_complex(cx)
        st = eng.run(100, 20, sink=lambda c, r: recs.append(r))
        state = eng.state(cid)
        recs = np.concatenate(recs)
        assert len(recs) == cx.n_rep + st["accepted"] == cx.n_rep + state["acs_mc"].sum()
        _ladders_are_permutations(cx, state)
        _records_are_exact(eng, cid, recs)
        rng = np.random.default_rng(0)
        pick = rng.choice(len(recs), 5000, replace=False)
        mm_r, ip_r, il_r = _random_poses(rng, cx, 5000)
        mm = np.concatenate([recs["movematrix"][pick], mm_r])
        ip = np.concatenate([recs["replica"][pick, 1], ip_r]).astype(np.int32)
        il = np.concatenate([recs["replica"][pick, 3], il_r]).astype(np.int32)
        out = eng.calc_energy(cid, il, ip, mm)
    worst = parity.check(out, _fp64(cx, ip, il, mm), label="ensemble")
    _say(f"10,000 poses vs FP64, worst (rel, err/mag): "
         f"{ {k: (f'{a:.1e}', f'{b:.1e}') for k, (a, b) in worst.items()} }; "
         f"{len(recs)} records exact; acceptance {st['accepted'] / (100 * cx.n_rep):.3f}")


def test_batch_256_full_size(gd):
    from oracle import parity
    cxs = synth.uneven_batch(256, seed=2)
    n_rep = sum(c.n_rep for c in cxs)
    _say(f"\nbatch: {len(cxs)} complexes, {n_rep} replicas")
    shape = dict(receptor_split=0, ladders_per_pack=0)
    states = {}
    for launch in (abi.LAUNCH_PERSISTENT, abi.LAUNCH_NO_GRAPH):
        recs = {c: [] for c in range(len(cxs))}
        with gd.Engine(seed=7, steps_per_dump=40, launch=launch, **shape) as eng:
            for cx in cxs:
                eng.add_complex(cx)
            eng.run(80, 20, sink=lambda c, r: recs[c].append(r))
            states[launch] = [eng.state(c) for c in range(len(cxs))]
            if launch != abi.LAUNCH_PERSISTENT:
                continue
            rng = np.random.default_rng(1)
            worst = {}
            for c, cx in enumerate(cxs):
                r = np.concatenate(recs[c])
                _ladders_are_permutations(cx, states[launch][c])
                _records_are_exact(eng, c, r)
                pick = rng.choice(len(r), 6, replace=False)
                mm_r, ip_r, il_r = _random_poses(rng, cx, 2)
                mm = np.concatenate([r["movematrix"][pick], mm_r])
                ip = np.concatenate([r["replica"][pick, 1], ip_r]).astype(np.int32)
                il = np.concatenate([r["replica"][pick, 3], il_r]).astype(np.int32)
                w = parity.check(eng.calc_energy(c, il, ip, mm), _fp64(cx, ip, il, mm),
                                 label=f"complex {c}")
                worst = {k: max(worst.get(k, (0, 0)), v) for k, v in w.items()}
    for a, b in zip(states[abi.LAUNCH_PERSISTENT], states[abi.LAUNCH_NO_GRAPH]):
        for k in a:
            assert np.array_equal(a[k], b[k]), k
    _say(f"2,048 poses vs FP64 over 256 complexes, worst (rel, err/mag): "
         f"{ {k: (f'{a:.1e}', f'{b:.1e}') for k, (a, b) in worst.items()} }")


def test_surrogate_training_2m_poses(gd):
    cxs = [synth.ensemble(seed=1)] + synth.uneven_batch(256, seed=2)
    with gd.Engine(seed=11, record=False, steps_per_dump=100, receptor_split=0,
                   ladders_per_pack=0, launch=abi.LAUNCH_PERSISTENT) as eng:
        for cx in cxs:
            eng.add_complex(cx)
        eng.run(40, 20)
        recs, rcid = sg.log_states(eng, 5, 20, 20)
        _say(f"\n{len(recs)} logged states from the ensemble and the 256-complex batch")
        eng.sur_build_grids(0.5, 14.0)
        ts = sg.build_training_set(eng, recs, rcid, 2_000_000, seed=1)
        m, sigma = sg.load_training_set(eng, ts)
        fit = sg.fit(eng, ts, sigma, n_train=int(0.9 * len(ts["t"])), epochs=20,
                     batch_poses=4096, lr=1e-3)
    _say(f"trained on {fit['n_train']} poses, 20 epochs in {fit['seconds']:.1f} s: held-out "
         f"RMSE {fit['test_rmse_energy']:.4f} (energy spread {fit['test_energy_std']:.4f}), loss {fit['history'][0]:.3f} -> "
         f"{fit['train_loss']:.3f}")
    assert fit["n_train"] >= 1_800_000 and fit["n_test"] >= 190_000
    assert np.isfinite(fit["test_rmse_energy"])
    # targets are standardized: predicting their mean scores a loss of about 1
    assert fit["train_loss"] <= 0.25
    assert fit["test_rmse_energy"] <= 0.4 * fit["test_energy_std"]


def test_da_1024_complexes(gd):
    cxs = synth.uneven_batch(1024, seed=2)
    n_rep = sum(c.n_rep for c in cxs)
    _say(f"\nDA batch: {len(cxs)} complexes, {n_rep} replicas")
    shape = dict(receptor_split=0, ladders_per_pack=0, launch=abi.LAUNCH_PERSISTENT)
    params = sg.pack_params(sg.init_params(3))
    finals = []
    for dump in (100, 40):
        recs = {c: [] for c in range(len(cxs))}
        with gd.Engine(seed=5, steps_per_dump=dump, **shape) as eng:
            for cx in cxs:
                eng.add_complex(cx)
            eng.sur_build_grids(0.8, 12.0)
            eng.sur_set_params(params)
            eng.sur_tabulate()
            st = eng.run_da(200, 20, 4, 2.0, sink=lambda c, r: recs[c].append(r))
            states = [eng.state(c) for c in range(len(cxs))]
            finals.append(states)
            if dump != 100:
                continue
            n_recs = 0
            for c, cx in enumerate(cxs):
                r = np.concatenate(recs[c])
                n_recs += len(r)
                s = states[c]
                _ladders_are_permutations(cx, s)
                assert np.all((r["step"][r["step"] > 0] + 1) % 4 == 0)
                if c % 8 == 0:
                    _records_are_exact(eng, c, r)
                    ex = eng.calc_energy(c, s["replica"][:, 3], s["replica"][:, 1],
                                         s["movematrix"])
                    assert np.array_equal(ex["e"].view(np.uint32),
                                          s["energy"]["e"].view(np.uint32))
            assert n_recs == n_rep + st["accepted"]
            _say(f"{n_recs} records; composite acceptance "
                 f"{st['accepted'] / (n_rep * 200 / 4):.3f}")
    for a, b in zip(*finals):
        for k in a:
            assert np.array_equal(a[k], b[k]), k
