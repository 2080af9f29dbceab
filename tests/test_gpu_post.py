"""Device k-means (gd_kmeans_medoids) and post_mc (gd_post_mc) against the
CPU oracle: identical medoids, cluster sizes and pass counts (the cluster
sums are exact integers, so no result depends on the order of device
atomics)."""
import numpy as np
import pytest

from geaxdock_gpu_amd import abi, synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gd():
    import geaxdock_gpu_amd as g
    return g


def _features(rng, n, dup=False):
    k_true = int(rng.integers(3, 12))
    centers = rng.normal(0, 3, size=(k_true, 9)) * rng.uniform(0.3, 6, size=9)
    lab = rng.integers(0, k_true, size=n)
    f = centers[lab] + rng.normal(0, 0.5, size=(n, 9))
    if dup and n > 10:
        f[n // 2:] = f[: n - n // 2]  # exact duplicates: ties everywhere
    return f.astype(np.float32)


@pytest.mark.parametrize("k", [50, 7])
def test_kmeans_groups_match_oracle(oracle, gd, k):
    rng = np.random.default_rng(k)
    sizes = [0, 5, k - 1, k, k + 1, 700, 3000, 1200, 2049]
    feats = [_features(rng, n, dup=(i == 6)) for i, n in enumerate(sizes)]
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    F = np.concatenate(feats)
    with gd.Engine() as eng:
        med, sz, ps = eng.kmeans_medoids(F, off, k, seed=11)
    for g, f in enumerate(feats):
        _, om, osz, op = oracle.kmeans(f, k, group=g, level=0, seed=11)
        assert np.array_equal(med[g], om), (g, sizes[g])
        assert np.array_equal(sz[g], osz), (g, sizes[g])
        assert ps[g] == op, (g, ps[g], op)
        assert sz[g].sum() == sizes[g]


def test_kmeans_one_large_group_k500(oracle, gd):
    """The shape of the second post_mc level: one group, 500 clusters."""
    rng = np.random.default_rng(3)
    f = _features(rng, 12000)
    with gd.Engine() as eng:
        med, sz, ps = eng.kmeans_medoids(f, [0, len(f)], 500, seed=2)
    _, om, osz, op = oracle.kmeans(f, 500, group=0, level=0, seed=2)
    assert np.array_equal(med[0], om) and np.array_equal(sz[0], osz) and ps[0] == op
    print(f"\n12,000 objects, k=500: {op} passes, {int((osz > 0).sum())} non-empty clusters")


def test_kmeans_rejects_bad_shapes(gd):
    f = np.zeros((10, 9), np.float32)
    with gd.Engine() as eng:
        with pytest.raises(gd.GeauxDockError):
            eng.kmeans_medoids(f, [0, 10], 513)
        with pytest.raises(gd.GeauxDockError):
            eng.kmeans_medoids(f, [0, 11, 10], 3)
        with pytest.raises(gd.GeauxDockError):
            eng.kmeans_medoids(f, [0, 10], 3, max_pass=0)


def test_post_mc_on_engine_records_matches_oracle(oracle, gd):
    """post_mc over the accepted-step records of a real engine run."""
    cx = synth.make_complex(4, lna=24, pnp=900, n_lig=4, n_prt=2, n_tmp=4)
    recs = []
    with gd.Engine(seed=3) as eng:
        eng.add_complex(cx)
        eng.run(600, 20, sink=lambda c, r: recs.append(r))
        allr = np.concatenate(recs)
        idx, sz = eng.post_mc(allr, 50, 500, seed=4)
    oidx, osz = oracle.post_mc(allr, 50, 500, seed=4)
    print(f"\n{len(allr)} records of {cx.n_rep} replicas -> {len(idx)} medoids")
    assert np.array_equal(idx, oidx) and np.array_equal(sz, osz)
    bound = sum(min(50, int((allr["replica"][:, 0] == r).sum())) for r in range(cx.n_rep))
    assert len(idx) > 0 and 0 < sz.sum() <= bound
