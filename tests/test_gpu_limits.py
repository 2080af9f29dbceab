"""The engine at the largest sizes the C ABI and the reference accept, and
just past them, through the C ABI.

* A 256-atom ligand (GD_MAX_LNA; 64 lanes x 4 atoms per replica): every
  energy term against the FP64 oracle, and an MC run's records equal to
  gd_calc_energy of their poses bit for bit.
* A 64-temperature ladder (GD_MAX_NTMP; 512 lanes): MC with exchanges keeps
  every ladder a permutation of its temperatures, records exact.
* A 10,000-point receptor with 10,000 KDE points (the reference's MAXPRO and
  MAXKDE, src/size.h:22,73), streamed through LDS tiles: terms against FP64
  and a short MC run with exact records.
* One past each limit is refused with GD_ERR_CAPACITY; zero poses and zero
  steps are ordinary calls.
"""
import numpy as np
import pytest

from geaxdock_gpu_amd import abi, synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gd():
    import geaxdock_gpu_amd as g
    return g


def _poses(rng, cx, n, t_sigma=1.0):
    mm = np.zeros((n, 6), np.float32)
    mm[:, :3] = rng.normal(0.0, t_sigma, (n, 3))
    mm[:, 3:] = rng.uniform(-np.pi, np.pi, (n, 3))
    mm[0] = 0.0
    return (rng.integers(0, cx.n_lig, n).astype(np.int32),
            rng.integers(0, cx.n_prt, n).astype(np.int32), mm)


def _fp64(cx, ip, il, mm):
    from oracle import gdo
    d = cx.desc()
    return gdo.energy_f64_batch(d, gdo.refmap_f64(d, cx), ip, il, mm)


def _run_exact(eng, cid, n_steps, nc):
    recs = []
    eng.run(n_steps, nc, sink=lambda c, r: recs.append(r))
    recs = np.concatenate(recs)
    ex = eng.calc_energy(cid, recs["replica"][:, 3], recs["replica"][:, 1], recs["movematrix"])
    assert np.array_equal(ex["e"].view(np.uint32), recs["energy"]["e"].view(np.uint32))
    return recs


def test_largest_ligand(gd):
    from oracle import parity
    cx = synth.make_complex(61, lna=abi.MAX_LNA, pnp=1500, box=40.0, pos=4, mcs_atoms=40,
                            n_lig=2, n_prt=1, n_tmp=4)
    rng = np.random.default_rng(3)
    il, ip, mm = _poses(rng, cx, 24)
    with gd.Engine(seed=2) as eng:
        cid = eng.add_complex(cx)
        assert eng.shape(cid) == (64, 4)
        out = eng.calc_energy(cid, il, ip, mm)
        parity.check(out, _fp64(cx, ip, il, mm), label="lna 256")
        recs = _run_exact(eng, cid, 40, 20)
    assert len(recs) > 0


def test_largest_ladder(gd):
    cx = synth.make_complex(62, lna=16, pnp=900, n_lig=2, n_prt=1, n_tmp=abi.MAX_NTMP)
    with gd.Engine(seed=3) as eng:
        cid = eng.add_complex(cx)
        lpr, apl = eng.shape(cid)
        assert lpr * cx.n_tmp == 512
        _run_exact(eng, cid, 80, 20)
        st = eng.state(cid)
    flat = np.arange(cx.n_rep)
    key = (flat // (cx.n_tmp * cx.n_lig)) * cx.n_lig + flat % cx.n_lig
    it = st["replica"][:, 2]
    for g in np.unique(key):
        assert np.array_equal(np.sort(it[key == g]), np.arange(cx.n_tmp))
    assert st["acs_x"].sum() > 0


def test_reference_maximum_receptor_and_kde(gd):
    from oracle import parity
    cx = synth.make_complex(63, lna=30, pnp=10000, pnk=10000, box=64.0, void_radius=8.0,
                            n_lig=2, n_prt=1, n_tmp=4)
    rng = np.random.default_rng(4)
    il, ip, mm = _poses(rng, cx, 16)
    with gd.Engine(seed=4) as eng:
        cid = eng.add_complex(cx)
        out = eng.calc_energy(cid, il, ip, mm)
        parity.check(out, _fp64(cx, ip, il, mm), label="pnp 10000, pnk 10000")
        _run_exact(eng, cid, 40, 20)


def test_one_past_each_limit_is_refused(gd):
    def status_of(cx):
        with gd.Engine(seed=1) as eng:
            with pytest.raises(gd.GeauxDockError) as ei:
                eng.add_complex(cx)
        return abi.STATUS[ei.value.status]

    assert status_of(synth.make_complex(64, lna=abi.MAX_LNA + 1, pnp=800, box=40.0, mcs_atoms=20,
                                        n_tmp=2)) == "GD_ERR_CAPACITY"
    assert status_of(synth.make_complex(65, lna=16, pnp=800, n_tmp=abi.MAX_NTMP + 1)) == \
        "GD_ERR_CAPACITY"


def test_zero_poses_and_zero_steps(gd):
    cx = synth.make_complex(66, lna=20, pnp=800, n_lig=1, n_prt=1, n_tmp=4)
    with gd.Engine(seed=5) as eng:
        cid = eng.add_complex(cx)
        out = eng.calc_energy(cid, np.zeros(0, np.int32), np.zeros(0, np.int32),
                              np.zeros((0, 6), np.float32))
        assert len(out) == 0
        recs = []
        st = eng.run(0, 20, sink=lambda c, r: recs.append(r))
        assert st["steps"] == 0
        s0 = eng.state(cid)
        # the start poses were evaluated and recorded; no move was made
        assert np.all(s0["movematrix"] == 0) and np.all(s0["acs_mc"] == 0)
        assert sum(len(r) for r in recs) == cx.n_rep


@pytest.mark.parametrize("pnp", [4200, 4600, 5600])
@pytest.mark.parametrize("split", [0, 8])
def test_receptors_near_the_lds_limit(gd, pnp, split):
    """Receptors around the size where the resident LDS image (receptor,
    type-pair records, the split's exchange buffers) stops fitting: each
    either stays resident or streams through tiles, at any split, with terms
    against FP64 and exact records."""
    from oracle import parity
    cx = synth.make_complex(80 + pnp // 100, lna=24, pnp=pnp, pnk=100, n_lig=2, n_prt=1,
                            n_tmp=4)
    rng = np.random.default_rng(pnp)
    il, ip, mm = _poses(rng, cx, 12)
    with gd.Engine(seed=6, receptor_split=split) as eng:
        cid = eng.add_complex(cx)
        out = eng.calc_energy(cid, il, ip, mm)
        parity.check(out, _fp64(cx, ip, il, mm), label=f"pnp {pnp} split {split}")
        _run_exact(eng, cid, 40, 20)
