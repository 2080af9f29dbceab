"""Post-MC clustering, trajectory output, the record file and the trace tool
on the CPU: the k-means oracle pinned against scikit-learn's Lloyd k-means
and brute force, and the library's host-side writers and readers.
"""
import os

import numpy as np
import pytest

from geaxdock_gpu_amd import GeauxDockError, abi, io

G = os.path.join(os.path.dirname(__file__), "golden", "inputs")
PARAS = os.path.join(G, "parameters", "paras")
A07 = [os.path.join(G, "1a07C1", f) for f in ("1a07C1.sdf", "1a07C.pdb", "1a07C1-0.8.ff")]


def blobs(rng, n, k_true=6, spread=0.4):
    centers = rng.normal(0, 3, size=(k_true, 9)) * rng.uniform(0.5, 4, size=9)
    lab = rng.integers(0, k_true, size=n)
    return (centers[lab] + rng.normal(0, spread, size=(n, 9))).astype(np.float32)


# ------------------------------------------------------------------ k-means oracle

@pytest.mark.parametrize("n,k,seed", [(3000, 50, 3), (500, 20, 1), (4000, 120, 7)])
def test_oracle_kmeans_matches_sklearn_lloyd(oracle, n, k, seed):
    """Run to full convergence (threshold 0), the oracle and scikit-learn's
    Lloyd k-means from the same initial centers give the same labels after
    the same number of passes."""
    from sklearn.cluster import KMeans
    rng = np.random.default_rng(seed)
    X = (rng.normal(size=(n, 9)) * rng.uniform(0.1, 5, size=9)).astype(np.float32)
    memb, med, sz, passes = oracle.kmeans(X, k, threshold=0.0, max_pass=100000, seed=seed)
    init = oracle.km_init(n, k, 0, 0, seed)
    km = KMeans(n_clusters=k, init=X[init].astype(np.float64), n_init=1, algorithm="lloyd",
                tol=0, max_iter=100000).fit(X.astype(np.float64))
    assert np.array_equal(km.labels_, memb)
    assert km.n_iter_ == passes
    assert sz.sum() == n and np.array_equal(sz, np.bincount(memb, minlength=k))


def test_oracle_medoid_is_member_nearest_its_mean(oracle):
    rng = np.random.default_rng(2)
    X = blobs(rng, 2000)
    memb, med, sz, _ = oracle.kmeans(X, 30, seed=5)
    for c in range(30):
        members = np.flatnonzero(memb == c)
        if len(members) == 0:
            assert med[c] == -1 and sz[c] == 0
            continue
        mean = X[members].astype(np.float64).mean(axis=0)
        d = ((X[members] - mean) ** 2).sum(axis=1)
        assert med[c] == members[np.argmin(d)]
        assert memb[med[c]] == c


def test_oracle_threshold_cap_and_small_groups(oracle):
    rng = np.random.default_rng(3)
    X = rng.normal(size=(800, 9)).astype(np.float32)
    _, _, _, p_full = oracle.kmeans(X, 40, threshold=0.0, max_pass=10000)
    _, _, _, p_thr = oracle.kmeans(X, 40, threshold=0.001)
    _, _, _, p_cap = oracle.kmeans(X, 40, threshold=0.0, max_pass=3)
    assert p_cap == 3 and p_thr <= p_full and p_full > 3
    # fewer objects than clusters: every object is its own medoid
    memb, med, sz, p = oracle.kmeans(X[:7], 10)
    assert p == 0 and list(med[:7]) == list(range(7)) and all(med[7:] == -1)
    assert list(sz) == [1] * 7 + [0] * 3


def test_km_init_is_first_k_of_a_uniform_permutation(oracle):
    n, k = 37, 5
    firsts = np.zeros(n, np.int64)
    for s in range(3700):
        init = oracle.km_init(n, k, 0, 0, s)
        assert len(set(init.tolist())) == k and init.min() >= 0 and init.max() < n
        firsts[init[0]] += 1
    from scipy.stats import chisquare
    assert chisquare(firsts).pvalue > 1e-4
    assert not np.array_equal(oracle.km_init(n, k, 0, 0, 1), oracle.km_init(n, k, 1, 0, 1))


def synth_records(rng, n_rep=12, per=(0, 400)):
    recs = []
    for r in range(n_rep):
        m = int(rng.integers(*per))
        x = np.zeros(m, abi.RECORD_DTYPE)
        x["replica"][:, 0] = r
        x["replica"][:, 1] = rng.integers(0, 3)
        x["replica"][:, 3] = rng.integers(0, 5)
        x["energy"]["e"][:, :9] = blobs(rng, m) if m else 0
        x["energy"]["e"][:, 9] = x["energy"]["e"][:, :9].sum(axis=1)
        x["movematrix"] = rng.normal(size=(m, 6))
        x["step"] = np.arange(m)
        recs.append(x)
    allr = np.concatenate(recs)
    return allr[rng.permutation(len(allr))]  # replicas interleaved


def test_oracle_post_mc_is_two_kmeans_levels(oracle):
    """post_mc = k-means of each replica's records (ascending replica id, the
    given order), then of all their medoids in replica and cluster order."""
    rng = np.random.default_rng(4)
    recs = synth_records(rng)
    got_idx, got_sz = oracle.post_mc(recs, 10, 25, seed=9)
    l1 = []
    g = 0
    for r in range(recs["replica"][:, 0].max() + 1):
        idx = np.flatnonzero(recs["replica"][:, 0] == r)
        if len(idx) == 0:
            continue
        _, med, _, _ = oracle.kmeans(recs["energy"]["e"][idx, :9], 10, group=g, level=1, seed=9)
        g += 1
        l1 += [idx[m] for m in med if m >= 0]
    l1 = np.array(l1)
    _, med2, sz2, _ = oracle.kmeans(recs["energy"]["e"][l1, :9], 25, group=0, level=2, seed=9)
    keep = med2 >= 0
    assert np.array_equal(got_idx, l1[med2[keep]])
    assert np.array_equal(got_sz, sz2[keep])
    assert got_sz.sum() == len(l1)


# ------------------------------------------------------------------ writers, files

def test_states_csv_matches_reference_format(tmp_path):
    rng = np.random.default_rng(5)
    recs = synth_records(rng, n_rep=3, per=(5, 9))
    recs["energy"]["cms"] = rng.uniform(0, 1, len(recs))
    recs["energy"]["rmsd"] = rng.uniform(0, 9, len(recs))
    p = str(tmp_path / "s.csv")
    io.write_states_csv(p, recs)
    lines = open(p).read().split("\n")
    assert lines[0] == ("lig prt cms rmsd mv1 mv2 mv3 mv4 mv5 mv6 "
                        "vdw ele pmf psp hdb hpc kde lhm edst ener")
    for r, l in zip(recs, lines[1:]):
        want = [str(r["replica"][3]), str(r["replica"][1]), "%.4f" % r["energy"]["cms"],
                "%.4f" % r["energy"]["rmsd"]] + ["%.4f" % v for v in r["movematrix"]] + \
               ["%.4f" % v for v in r["energy"]["e"]]
        assert l == " ".join(want)
    assert lines[len(recs) + 1] == "" and len(lines) == len(recs) + 2


def test_medoids_csv_matches_reference_format(tmp_path):
    rng = np.random.default_rng(6)
    recs = synth_records(rng, n_rep=2, per=(3, 5))
    sz = rng.integers(1, 40, len(recs)).astype(np.int32)
    p = str(tmp_path / "m.csv")
    io.write_medoids_csv(p, recs, sz)
    lines = open(p).read().split("\n")
    assert lines[0].split() == ("lig prt t0 t1 t2 r0 r1 r2 vdw ele pmf psp hdb hpc kde lhm "
                                "edst cms rmsd ener cluster_sz").split()
    for r, s, l in zip(recs, sz, lines[1:]):
        want = [str(r["replica"][3]), str(r["replica"][1])] + \
               ["%f" % v for v in r["movematrix"]] + ["%f" % v for v in r["energy"]["e"][:9]] + \
               ["%f" % r["energy"]["cms"], "%f" % r["energy"]["rmsd"], "%f" % r["energy"]["e"][9],
                str(s)]
        assert l == " ".join(want) + " "


def test_record_file_round_trip(tmp_path):
    rng = np.random.default_rng(7)
    a = synth_records(rng, n_rep=4, per=(10, 30))
    b = synth_records(rng, n_rep=2, per=(5, 8))
    p = str(tmp_path / "r.gdrec")
    f = io.RecordFile(p)
    assert len(f) == 0
    f.write(a)
    f.sink(0, b)
    assert len(f) == len(a) + len(b)
    assert os.path.getsize(p) == 64 + 92 * (len(a) + len(b))
    allr = f.read()
    assert allr.tobytes() == np.concatenate([a, b]).tobytes()
    assert f.read(len(a), len(b)).tobytes() == b.tobytes()
    assert np.asarray(f.memmap()).tobytes() == allr.tobytes()
    # reopen for appending keeps what is there
    g = io.RecordFile(p, append=True)
    g.write(a[:3])
    assert len(g) == len(a) + len(b) + 3
    with pytest.raises(GeauxDockError, match="past"):
        g.read(len(g) - 1, 2)
    bad = str(tmp_path / "bad")
    open(bad, "wb").write(b"x" * 100)
    with pytest.raises(GeauxDockError, match="not a record file"):
        io.RecordFile(bad, append=True).read()


# ------------------------------------------------------------------ trace tool

@pytest.fixture(scope="module")
def a07():
    return io.load_complex(*A07, PARAS, complex_id="1a07C1")


def test_pose_coords_match_oracle_placement(oracle, a07):
    d = a07.desc()
    rng = np.random.default_rng(8)
    for i in range(20):
        il = int(rng.integers(0, a07.n_lig))
        mm = np.concatenate([rng.normal(0, 2, 3), rng.uniform(-4, 4, 3)]).astype(np.float32)
        got = io.pose_coords(a07, il, mm)
        x, y, z = oracle.place_f64(d, a07, il, mm)
        np.testing.assert_allclose(got, np.stack([x, y, z], axis=1), atol=2e-5)
        # rigid: the centroid is the pocket centre plus the translation
        np.testing.assert_allclose(got.mean(axis=0), a07.pocket_center + mm[:3], atol=2e-5)


def test_trace_tool_writes_placed_sdf(tmp_path, a07):
    """The reference's own trace input (data/parameters/trace.txt) and three
    poses of data/trajectory.out: each written SDF keeps every line but the
    first record's atom coordinates, which are the placed pose."""
    rows = io.read_trace(os.path.join(G, "parameters", "trace.txt"))
    assert len(rows) == 2 and rows[1][2] == pytest.approx(0.89)
    traj = io.read_trace(os.path.join(G, "trajectory.out"))[:3]
    paths = io.trace(a07, A07[0], rows + traj, str(tmp_path / "pose"))
    src = open(A07[0]).read().split("\n")
    for r, p in zip(rows + traj, paths):
        out = open(p).read().split("\n")
        assert len(out) == len(src) + (0 if src[-1] == "" else 1)
        want = io.pose_coords(a07, int(r[1]), r[2:8])
        for a in range(a07.lna):
            l = out[4 + a]
            xyz = [float(l[0:10]), float(l[10:20]), float(l[20:30])]
            np.testing.assert_allclose(xyz, want[a], atol=5.1e-5)
            assert l[30:] == src[4 + a][30:]
        assert out[:4] == src[:4] and out[4 + a07.lna:len(src)] == src[4 + a07.lna:]
    # the zero move of ligand conformer 0 puts its centroid on the pocket centre
    back = io.sdf_read(paths[0])
    np.testing.assert_allclose(back.xyz[0].mean(axis=1), a07.pocket_center, atol=1e-4)
