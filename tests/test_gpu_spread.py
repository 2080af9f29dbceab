"""Spread packing (GD_LAUNCH_SPREAD) on the GPU: one wave of replicas per
workgroup at a receptor split of 8, a ladder over several workgroups, and the
replica exchanges in a kernel of their own between launches.

* Ladders of 3 to 40 temperatures (several packs each) and a single complex
  follow the FP32 CPU path decision for decision, exchanges included, up to
  explained rounding splits (oracle/decisions.py); records carry the bits
  gd_calc_energy gives their poses; graph-captured, eager and persistent
  launches agree bit for bit.
* A run is a function of the seed alone: record windows that do not line up
  with the exchange interval give the same states and records.
* Delayed acceptance with surrogate scale 0 follows the CPU DA chain.
* Spread packing refuses a fixed receptor split or ladders per pack; it is
  the default for workloads of at most one ladder per two CUs.
"""
import numpy as np
import pytest

from geaxdock_gpu_amd import abi, synth
from geaxdock_gpu_amd import surrogate as sg

pytestmark = pytest.mark.gpu

SPREAD = abi.LAUNCH_SPREAD


@pytest.fixture(scope="module")
def gd():
    import geaxdock_gpu_amd as g
    return g


def _cpu(oracle, cx, seed, cid, n_steps, nc, da_k=0):
    d = cx.desc()
    cpu = oracle.McState(cx, rec_cap=n_steps + 1)
    cpu.enable_log(0, n_steps, n_steps // nc)
    oracle.mc_run(d, cpu, seed=seed, cid=cid, init=True, step0=0, n_steps=0, nc=nc)
    if da_k:
        oracle.mc_run_da(d, cpu, seed=seed, cid=cid, step0=0, n_steps=n_steps, nc=nc, k_sub=da_k)
    else:
        oracle.mc_run(d, cpu, seed=seed, cid=cid, init=False, step0=0, n_steps=n_steps, nc=nc)
    return cpu


def test_spread_ladders_track_cpu_path(gd):
    from oracle import decisions, gdo
    shapes = [(3, 20, 5), (6, 40, 3), (12, 36, 3), (20, 14, 2), (40, 12, 1)]
    cxs = [synth.make_complex(50 + i, lna=lna, pnp=700 + 150 * i, n_lig=n_lig, n_prt=1,
                              n_tmp=n_tmp)
           for i, (n_tmp, lna, n_lig) in enumerate(shapes)]
    cxs.append(synth.single_complex(seed=0))
    n_steps, nc = 120, 20
    states = {}
    for launch in (SPREAD, SPREAD | abi.LAUNCH_PERSISTENT, SPREAD | abi.LAUNCH_NO_GRAPH):
        recs = {c: [] for c in range(len(cxs))}
        with gd.Engine(seed=9, launch=launch, steps_per_dump=60) as eng:
            for cx in cxs:
                eng.add_complex(cx)
            st = eng.run(n_steps, nc, sink=lambda c, r: recs[c].append(r))
            states[launch] = [eng.state(c) for c in range(len(cxs))]
            assert st["accepted"] == sum(int(s["acs_mc"].sum()) for s in states[launch])
            if launch != SPREAD:
                continue
            for c, cx in enumerate(cxs):
                r = np.concatenate(recs[c])
                ex = eng.calc_energy(c, r["replica"][:, 3], r["replica"][:, 1], r["movematrix"])
                assert np.array_equal(ex["e"].view(np.uint32), r["energy"]["e"].view(np.uint32))
                cpu = _cpu(gdo, cx, 9, c, n_steps, nc)
                out = decisions.compare(cx, r, cpu, nc, gpu_state=states[launch][c])
                print(f"\ncomplex {c} (n_tmp {cx.n_tmp}, lna {cx.lna}): "
                      f"{out['identical']}/{cx.n_rep} identical, explained {out['explained']}",
                      flush=True)
                flat = np.arange(cx.n_rep)
                key = (flat // (cx.n_tmp * cx.n_lig)) * cx.n_lig + flat % cx.n_lig
                it = states[launch][c]["replica"][:, 2]
                for g in np.unique(key):
                    assert np.array_equal(np.sort(it[key == g]), np.arange(cx.n_tmp))
    ref = states[SPREAD]
    for launch, st in states.items():
        for a, b in zip(ref, st):
            for k in a:
                assert np.array_equal(a[k], b[k]), (launch, k)


def test_spread_windows_off_the_exchange_grid(gd):
    """Record windows of 50 steps (exchanges every 20: sub-launches of 20, 20,
    10, then 10, 20, 20, ...) and one window of 300 give the same states and
    the same records."""
    cx = synth.make_complex(71, lna=24, pnp=900, n_lig=3, n_prt=2, n_tmp=8)

    def go(dump):
        recs = []
        with gd.Engine(seed=21, launch=SPREAD, steps_per_dump=dump) as eng:
            cid = eng.add_complex(cx)
            eng.run(300, 20, sink=lambda c, r: recs.append(r))
            st = eng.state(cid)
        r = np.concatenate(recs)
        return st, r[np.lexsort((r["step"], r["replica"][:, 0]))]

    s1, r1 = go(50)
    s2, r2 = go(300)
    for k in s1:
        assert np.array_equal(s1[k], s2[k]), k
    assert len(r1) > cx.n_rep and r1.tobytes() == r2.tobytes()
    assert s1["acs_x"].sum() > 0


def test_spread_da_scale0_tracks_cpu_path(gd):
    from oracle import decisions, gdo
    cx = synth.single_complex(seed=0)
    n_steps, nc, k = 200, 20, 4
    recs = []
    with gd.Engine(seed=11, launch=SPREAD) as eng:
        cid = eng.add_complex(cx)
        eng.sur_build_grids(0.8, 12.0)
        eng.sur_set_params(sg.pack_params(sg.init_params(3)))
        eng.run_da(n_steps, nc, k, 0.0, sink=lambda c, r: recs.append(r))
        state = eng.state(cid)
    recs = np.concatenate(recs)
    cpu = _cpu(gdo, cx, 11, 0, n_steps, nc, da_k=k)
    out = decisions.compare(cx, recs, cpu, nc, k_sub=k, gpu_state=state, da=True)
    print("\nspread DA(scale 0):", out["identical"], "of", cx.n_rep, "identical, explained",
          out["explained"])


def test_spread_refuses_fixed_packing(gd):
    cx = synth.single_complex(seed=3)
    for kw in (dict(receptor_split=2), dict(ladders_per_pack=1)):
        with gd.Engine(seed=1, launch=SPREAD, **kw) as eng:
            eng.add_complex(cx)
            with pytest.raises(gd.GeauxDockError) as ei:
                eng.run(20, 20)
            assert abi.STATUS[ei.value.status] == "GD_ERR_INVALID"


def test_spread_is_automatic_for_small_workloads(gd):
    """A workload of at most one ladder per two CUs gets spread packing by
    default (the same bits as asking for it); NO_SPREAD keeps ladders whole,
    with its own bits, and records still equal gd_calc_energy of their poses."""
    cx = synth.make_complex(72, lna=28, pnp=1000, n_lig=3, n_prt=1, n_tmp=8)

    def go(launch):
        recs = []
        with gd.Engine(seed=5, launch=launch) as eng:
            cid = eng.add_complex(cx)
            eng.run(100, 20, sink=lambda c, r: recs.append(r))
            st = eng.state(cid)
            r = np.concatenate(recs)
            ex = eng.calc_energy(cid, r["replica"][:, 3], r["replica"][:, 1], r["movematrix"])
            assert np.array_equal(ex["e"].view(np.uint32), r["energy"]["e"].view(np.uint32))
        return st

    auto, spread, whole = go(0), go(SPREAD), go(abi.LAUNCH_NO_SPREAD)
    for k in auto:
        assert np.array_equal(auto[k], spread[k]), k
    assert not np.array_equal(auto["energy"]["e"].view(np.uint32),
                              whole["energy"]["e"].view(np.uint32))


def test_spread_widens_lane_shapes_once(gd):
    """Under spread packing a complex that goes to the device for the first
    time takes the lane shape that makes the most useful packs (the single
    complex: 32 lanes of one atom instead of 16 lanes of two); the shape then
    stays, so energies before and after a run have the same bits. A forced
    shape, or NO_SPREAD, keeps the shape add_complex chose."""
    cx = synth.single_complex(seed=4)
    rng = np.random.default_rng(8)
    mm = np.zeros((6, 6), np.float32)
    mm[:, :3] = rng.normal(0, 0.5, (6, 3))
    il = np.zeros(6, np.int32)
    with gd.Engine(seed=2) as eng:
        cid = eng.add_complex(cx)
        assert eng.shape(cid) == (16, 2)
        e0 = eng.calc_energy(cid, il, il, mm)
        assert eng.shape(cid) == (32, 1)
        recs = []
        eng.run(60, 20, sink=lambda c, r: recs.append(r))
        assert eng.shape(cid) == (32, 1)
        e1 = eng.calc_energy(cid, il, il, mm)
        assert np.array_equal(e0["e"].view(np.uint32), e1["e"].view(np.uint32))
        r = np.concatenate(recs)
        ex = eng.calc_energy(cid, r["replica"][:, 3], r["replica"][:, 1], r["movematrix"])
        assert np.array_equal(ex["e"].view(np.uint32), r["energy"]["e"].view(np.uint32))
    for kw in (dict(lanes_per_replica=16), dict(launch=abi.LAUNCH_NO_SPREAD)):
        with gd.Engine(seed=2, **kw) as eng:
            cid = eng.add_complex(cx)
            eng.calc_energy(cid, il, il, mm)
            assert eng.shape(cid) == (16, 2), kw
