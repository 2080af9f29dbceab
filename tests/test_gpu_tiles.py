"""Receptors streamed through LDS tiles (receptors and KDE sets beyond LDS).

The reference accepts up to 10,000 receptor points and 10,000 KDE points
(src/size.h:22,73) and reads every receptor point from global memory in its
pair loop (src/kernel_cuda_l2_calcenergy.cu:92-100). Here a receptor that
does not fit in LDS streams through double-buffered LDS tiles and its KDE
points are read from HBM; gd_opts.receptor_tile forces tiles of a given
length on any receptor. Checked against the FP64 oracle term by term, against
the FP32 CPU path decision by decision, and for self-consistency (records
equal gd_calc_energy of their poses bit for bit, launch modes bit-identical).
"""
import numpy as np
import pytest

from geaxdock_gpu_amd import abi, synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gd():
    import geaxdock_gpu_amd as g
    return g


def _poses(rng, cx, n, t_sigma=1.5):
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


def _mc_vs_cpu(gd, cx, seed, n_steps, nc, **opts):
    from oracle import decisions, gdo
    recs = []
    with gd.Engine(seed=seed, **opts) as eng:
        cid = eng.add_complex(cx)
        eng.run(n_steps, nc, sink=lambda c, r: recs.append(r))
        state = eng.state(cid)
        recs = np.concatenate(recs)
        ex = eng.calc_energy(cid, recs["replica"][:, 3], recs["replica"][:, 1],
                             recs["movematrix"])
        assert np.array_equal(ex["e"].view(np.uint32), recs["energy"]["e"].view(np.uint32))
    d = cx.desc()
    cpu = gdo.McState(cx, rec_cap=n_steps + 1)
    cpu.enable_log(0, n_steps, n_steps // nc)
    gdo.mc_run(d, cpu, seed=seed, init=True, step0=0, n_steps=0, nc=nc)
    gdo.mc_run(d, cpu, seed=seed, init=False, step0=0, n_steps=n_steps, nc=nc)
    return decisions.compare(cx, recs, cpu, nc, gpu_state=state)


def test_receptor_and_kde_beyond_lds(gd):
    """8,000 receptor points and 4,000 KDE points (about 250 KB resident, more
    than the 160 KB of LDS): the engine streams the receptor in tiles."""
    from oracle import parity
    cx = synth.make_complex(21, lna=48, pnp=8000, pnk=4000, pos=6, mcs_atoms=30, n_lig=2,
                            n_prt=2, n_tmp=4)
    rng = np.random.default_rng(2)
    il, ip, mm = _poses(rng, cx, 200)
    with gd.Engine(seed=1) as eng:
        cid = eng.add_complex(cx)
        out = eng.calc_energy(cid, il, ip, mm)
    worst = parity.check(out, _fp64(cx, ip, il, mm), label="pnp 8000, pnk 4000")
    print("\nworst (rel, err/mag):", {k: (f"{a:.1e}", f"{b:.1e}") for k, (a, b) in worst.items()})
    out = _mc_vs_cpu(gd, cx, 9, 60, 20)
    print("decisions:", out["identical"], "of", cx.n_rep, "identical; explained",
          out["explained"])


@pytest.mark.parametrize("tile", [64, 200, 512, 1024])
def test_forced_tiles_match_fp64(gd, tile):
    """Tiles of 64 to 1,024 points on a receptor that would fit in LDS: the
    same terms within the FP64 bound, and within 1e-5 of the resident
    layout's values (only the order of the chunk sums changes)."""
    from oracle import parity
    cx = synth.make_complex(5, lna=40, pnp=1500, pnk=150, pos=6, mcs_atoms=30, n_lig=3, n_prt=2,
                            n_tmp=2, void_radius=8.0)
    rng = np.random.default_rng(5)
    il, ip, mm = _poses(rng, cx, 64, t_sigma=2.0)
    outs = {}
    for t in (0, tile):
        with gd.Engine(seed=2, receptor_tile=t) as eng:
            cid = eng.add_complex(cx)
            outs[t] = eng.calc_energy(cid, il, ip, mm)
    parity.check(outs[tile], _fp64(cx, ip, il, mm), label=f"tile {tile}")
    np.testing.assert_allclose(outs[tile]["e"], outs[0]["e"], rtol=1e-5, atol=1e-5)
    assert np.array_equal(outs[tile]["cms"], outs[0]["cms"])


@pytest.mark.parametrize("tile,split", [(128, 1), (256, 2)])
def test_tiled_mc_tracks_cpu_path(gd, tile, split):
    cx = synth.single_complex(seed=0)
    out = _mc_vs_cpu(gd, cx, 7, 200, 20, receptor_tile=tile, receptor_split=split)
    print(f"\ntile {tile} split {split}: {out['identical']}/{cx.n_rep} identical, "
          f"explained {out['explained']}")


def test_tiled_batch_launch_modes_and_da(gd):
    """A batch mixing resident and tiled receptors (two launch classes per
    lane shape): graph-captured persistent and eager launches bit-identical;
    delayed acceptance on it keeps exact energies in its states."""
    from geaxdock_gpu_amd import surrogate as sg
    cxs = synth.uneven_batch(3, seed=5, lna_range=(12, 40), pnp_range=(800, 1600))
    cxs.append(synth.make_complex(31, lna=30, pnp=7000, pnk=3000, n_lig=2, n_prt=2, n_tmp=8))
    ref = None
    for launch in (abi.LAUNCH_PERSISTENT, abi.LAUNCH_NO_GRAPH):
        with gd.Engine(seed=3, record=False, launch=launch, steps_per_dump=40) as eng:
            for cx in cxs:
                eng.add_complex(cx)
            eng.run(80, 20)
            st = [eng.state(c) for c in range(len(cxs))]
        if ref is None:
            ref = st
            continue
        for a, b in zip(ref, st):
            for k in a:
                assert np.array_equal(a[k], b[k]), (launch, k)
    with gd.Engine(seed=4, launch=abi.LAUNCH_PERSISTENT) as eng:
        for cx in cxs:
            eng.add_complex(cx)
        eng.sur_build_grids(0.8, 12.0)
        eng.sur_set_params(sg.pack_params(sg.init_params(3)))
        eng.sur_tabulate()
        eng.run_da(80, 20, 4, 2.0)
        for c in range(len(cxs)):
            s = eng.state(c)
            ex = eng.calc_energy(c, s["replica"][:, 3], s["replica"][:, 1], s["movematrix"])
            assert np.array_equal(ex["e"].view(np.uint32), s["energy"]["e"].view(np.uint32)), c
