"""gd_calc_energy_noalloc (no device allocation; pose tables and energies in
the library's file-scope device scratch) against gd_calc_energy, through the
C ABI: bit-identical energies on two calls in a row, the second smaller (the
scratch keeps the first call's contents), and clear errors beyond the pose
limit and before the complex is on the device."""
import numpy as np
import pytest

from geaxdock_gpu_amd import abi, synth

pytestmark = pytest.mark.gpu


def _poses(rng, cx, n):
    mm = np.zeros((n, 6), np.float32)
    mm[:, :3] = rng.normal(0.0, 1.5, (n, 3))
    mm[:, 3:] = rng.uniform(-np.pi, np.pi, (n, 3))
    return (rng.integers(0, cx.n_lig, n).astype(np.int32),
            rng.integers(0, cx.n_prt, n).astype(np.int32), mm)


def test_noalloc_matches_calc_energy_twice_second_smaller():
    import geaxdock_gpu_amd as gd
    cx = synth.make_complex(5, lna=40, pnp=1500, n_lig=3, n_prt=4, n_tmp=2)
    rng = np.random.default_rng(0)
    with gd.Engine(seed=1) as eng:
        cid = eng.add_complex(cx)
        with pytest.raises(gd.GeauxDockError) as ei:
            eng.calc_energy_noalloc(cid, *_poses(rng, cx, 4))
        assert abi.STATUS[ei.value.status] == "GD_ERR_STATE"
        eng.calc_energy(cid, *_poses(rng, cx, 1))  # puts the complex on the device
        for n in (20000, 777):
            il, ip, mm = _poses(rng, cx, n)
            ref = eng.calc_energy(cid, il, ip, mm)
            got = eng.calc_energy_noalloc(cid, il, ip, mm)
            assert np.array_equal(got.view(np.uint8), ref.view(np.uint8)), n
        il, ip, mm = _poses(rng, cx, abi.NOALLOC_MAX_POSES + 1)
        with pytest.raises(gd.GeauxDockError) as ei:
            eng.calc_energy_noalloc(cid, il, ip, mm)
        assert abi.STATUS[ei.value.status] == "GD_ERR_CAPACITY" and "poses" in str(ei.value)
