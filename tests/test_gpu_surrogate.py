"""GPU surrogate (BF16 MFMA MLP) against the FP64 oracle (oracle/surrogate.py).

Features: grid-interpolated radial features vs the FP64 grid + trilinear
oracle, |err| <= 1e-3 + 1e-3 |f| (FP32 grid and interpolation).
Loss and gradients of one batch vs FP64 backprop on the same features:
BF16 operands (8-bit mantissa) bound the agreement, so each parameter group
must have cosine >= 0.999 and relative norm error <= 3e-2, and the loss
relative error <= 2e-2.
"""
import numpy as np
import pytest

from geaxdock_gpu_amd import synth
from oracle import surrogate as so

pytestmark = pytest.mark.gpu


def to_gpu_params(p):
    v = np.zeros(64 * 128 + 128 * 128 + 128 + 128 + 1, np.float32)
    W1 = np.zeros((64, 128))
    W1[:56] = p["W1"]
    W1[56] = p["b1"]
    o = 0
    v[o:o + 8192] = W1.reshape(-1)
    o += 8192
    v[o:o + 16384] = p["W2"].reshape(-1)
    o += 16384
    v[o:o + 128] = p["b2"]
    o += 128
    v[o:o + 128] = p["w3"]
    o += 128
    v[o] = p["b3"][0]
    return v


def from_gpu_grad(g):
    W1 = g[:8192].reshape(64, 128)
    return {"W1": W1[:56], "b1": W1[56], "W2": g[8192:8192 + 16384].reshape(128, 128),
            "b2": g[24576:24704], "w3": g[24704:24832], "b3": g[24832:24833]}


def poses(rng, n, sigma=2.0):
    mm = np.zeros((n, 6), np.float32)
    mm[:, :3] = rng.normal(0, sigma, (n, 3))
    mm[:, 3:] = rng.uniform(-np.pi, np.pi, (n, 3))
    return mm


@pytest.fixture(scope="module")
def setup():
    import geaxdock_gpu_amd as gd
    cx = synth.single_complex(seed=2, void_radius=7.0)
    eng = gd.Engine(seed=1)
    cid = eng.add_complex(cx)
    spacing, half = 0.8, 9.6
    eng.sur_build_grids(spacing, half)
    yield eng, cid, cx, spacing, half
    eng.close()


def test_features_match_fp64_grid(setup):
    eng, cid, cx, spacing, half = setup
    rng = np.random.default_rng(0)
    mm = poses(rng, 6)
    feat = eng.sur_features(cid, np.zeros(6), np.zeros(6), mm)
    n = int(np.ceil(2 * half / spacing)) + 1
    origin = np.float64(cx.pocket_center) - half
    prt = np.stack([cx.prt_x, cx.prt_y, cx.prt_z], 1)[:cx.pnp]
    grid = so.grid_features(origin, spacing, n, prt, cx.prt_t)
    from oracle import gdo
    d = cx.desc()
    for i in range(6):
        xyz = gdo.place_f64(d, cx, 0, mm[i]).T
        want = so.features(cx.lig_t, so.interp(grid, origin, spacing, xyz))
        got = feat[i]
        assert np.all(np.abs(got - want) <= 1e-3 + 1e-3 * np.abs(want)), np.max(np.abs(got - want))
    # grid interpolation approximates the direct features
    direct = so.radial_features(xyz, prt, cx.prt_t)
    rel = np.linalg.norm(feat[5][:, 24:] - direct) / np.linalg.norm(direct)
    assert rel < 0.05, rel


def test_loss_and_gradients_match_fp64(setup):
    eng, cid, cx, _, _ = setup
    rng = np.random.default_rng(1)
    n = 96
    mm = poses(rng, n)
    target = rng.normal(0, 2, n).astype(np.float32)
    eng.sur_dataset_clear()
    eng.sur_dataset_add(cid, np.zeros(n), np.zeros(n), mm, target)
    p = so.init_params(3)
    eng.sur_set_params(to_gpu_params(p))
    loss, g = eng.sur_loss_grad(0, n)
    F = eng.sur_features(cid, np.zeros(n), np.zeros(n), mm).reshape(-1, 56).astype(np.float64)
    pose = np.repeat(np.arange(n), cx.lna)
    lo, go = so.loss_and_grads(p, F, pose, target.astype(np.float64))
    assert abs(loss - lo) <= 2e-2 * lo, (loss, lo)
    gg = from_gpu_grad(g)
    for k in go:
        a, b = gg[k].reshape(-1).astype(np.float64), go[k].reshape(-1)
        cos = a @ b / (np.linalg.norm(a) * np.linalg.norm(b) + 1e-300)
        err = np.linalg.norm(a - b) / (np.linalg.norm(b) + 1e-300)
        print(k, f"cos {cos:.6f} rel {err:.2e}")
        assert cos >= 0.999 and err <= 3e-2, (k, cos, err)
    # batches split anywhere give the same gradient sum (row padding is inert)
    l1, g1 = eng.sur_loss_grad(0, 37)
    l2, g2 = eng.sur_loss_grad(37, n)
    assert np.isfinite(l1) and np.isfinite(l2) and np.all(np.isfinite(g1 + g2))


def test_training_reduces_loss(setup):
    eng, cid, cx, _, _ = setup
    rng = np.random.default_rng(2)
    n = 1024
    mm = poses(rng, n)
    # target: a smooth function of the pose the features can represent
    from oracle import gdo
    d = cx.desc()
    ref = gdo.refmap_f64(d, cx)
    target = np.array([gdo.energy_f64(d, ref, 0, 0, m).e[0] for m in mm], np.float32)
    target = (target - target.mean()) / target.std()
    eng.sur_dataset_clear()
    eng.sur_dataset_add(cid, np.zeros(n), np.zeros(n), mm, target)
    eng.sur_set_params(to_gpu_params(so.init_params(5)))
    l0, _ = eng.sur_loss_grad(0, n, grad=False)
    hist, secs = eng.sur_train(0, n, epochs=30, batch_poses=128, lr=2e-3)
    print("loss", l0, "->", hist[-1], f"{secs:.2f} s")
    assert hist[-1] < 0.35 * l0
    pred = eng.sur_predict(0, n)
    assert np.mean((pred - target) ** 2) == pytest.approx(hist[-1], rel=1e-3)


def test_dataset_add_multi_matches_per_complex_adds():
    """Poses of several complexes appended in one interleaved call give the
    same rows (hence predictions) as per-complex appends."""
    import geaxdock_gpu_amd as gd
    cxs = synth.uneven_batch(2, seed=8, lna_range=(12, 30), pnp_range=(800, 1200))
    rng = np.random.default_rng(3)
    n = 40
    cid = rng.integers(0, 2, n).astype(np.int32)
    il = np.array([rng.integers(0, cxs[c].n_lig) for c in cid], np.int32)
    ip = np.array([rng.integers(0, cxs[c].n_prt) for c in cid], np.int32)
    mm = poses(rng, n, sigma=1.0)
    t = rng.normal(0, 1, n).astype(np.float32)
    preds = []
    for multi in (True, False):
        with gd.Engine(seed=1) as eng:
            for cx in cxs:
                eng.add_complex(cx)
            eng.sur_build_grids(0.8, 9.6)
            eng.sur_set_params(to_gpu_params(so.init_params(2)))
            if multi:
                eng.sur_dataset_add_multi(cid, il, ip, mm, t)
                preds.append(eng.sur_predict(0, n))
            else:
                order = np.concatenate([np.nonzero(cid == c)[0] for c in (0, 1)])
                for c in (0, 1):
                    sel = np.nonzero(cid == c)[0]
                    eng.sur_dataset_add(c, il[sel], ip[sel], mm[sel], t[sel])
                p = np.zeros(n, np.float32)
                p[order] = eng.sur_predict(0, n)
                preds.append(p)
    np.testing.assert_array_equal(preds[0], preds[1])
