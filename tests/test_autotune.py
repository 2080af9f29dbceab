"""The autotuner's search space (no GPU): shape grid, tile stage, signature."""
from geaxdock_gpu_amd import abi, autotune, synth


def test_shape_grid_and_tile_stage():
    cxs = synth.uneven_batch(2, seed=1, lna_range=(12, 20), pnp_range=(800, 900))
    grid = autotune.candidates(cxs)
    assert len(grid) == 3 * 5 * 2
    assert {c["receptor_split"] for c in grid} == {0, 1, 2}
    assert {c["launch"] for c in grid} == {0, abi.LAUNCH_PERSISTENT}
    assert all(c["receptor_tile"] == 0 for c in grid)
    assert len(autotune.candidates(cxs, full=True)) == len(grid) + 3
    best = grid[7]
    tiles = autotune.tile_candidates(best)
    assert [t["receptor_tile"] for t in tiles] == [64, 128, 256, 512]
    for t in tiles:
        assert {k: v for k, v in t.items() if k != "receptor_tile"} == \
            {k: v for k, v in best.items() if k != "receptor_tile"}


def test_signature_depends_on_shapes_only():
    a = synth.uneven_batch(2, seed=1, lna_range=(12, 20), pnp_range=(800, 900))
    b = synth.uneven_batch(2, seed=1, lna_range=(12, 20), pnp_range=(800, 900))
    assert autotune.signature(a, 20) == autotune.signature(b, 20)
    assert autotune.signature(a, 20) != autotune.signature(a, 10)
