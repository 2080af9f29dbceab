"""The launch-shape autotuner on the GPU: both stages run (shapes, then
receptor tile lengths), every accepted candidate gives a time, and the
returned options are the fastest trial and drive an engine."""
import numpy as np
import pytest

from geaxdock_gpu_amd import abi, autotune, synth

pytestmark = pytest.mark.gpu


def test_autotune_shapes_then_tiles():
    import geaxdock_gpu_amd as gd
    cxs = synth.uneven_batch(3, seed=5, lna_range=(12, 40), pnp_range=(800, 1600))
    shapes = [dict(receptor_split=s, ladders_per_pack=0, launch=l, lanes_per_replica=0,
                   receptor_tile=0) for s in (0, 1, 2) for l in (0, abi.LAUNCH_PERSISTENT)]
    best, trials = autotune.tune(cxs, steps=8, warmup=4, budget_s=300, shapes=shapes)
    assert len(trials) == len(shapes) + len(autotune.TILES)
    tiles = [t["opts"]["receptor_tile"] for t in trials[len(shapes):]]
    assert tiles == list(autotune.TILES)
    ok = [t for t in trials if t["s_per_step"] is not None]
    assert len(ok) == len(trials)
    assert best == min(ok, key=lambda t: t["s_per_step"])["opts"]
    print("\n" + "\n".join(f"{t['opts']}: {t['s_per_step'] * 1e3:.3f} ms/step" for t in trials))
    with gd.Engine(seed=1, record=False, **best) as eng:
        for cx in cxs:
            eng.add_complex(cx)
        st = eng.run(20, 20)
    assert st["accepted"] > 0
