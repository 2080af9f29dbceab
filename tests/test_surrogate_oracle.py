"""Pins the FP64 surrogate oracle (oracle/surrogate.py): finite-difference
gradients, feature continuity at the cutoff, exact interpolation on grid nodes
and of linear fields, one-hot layout."""
import numpy as np

from oracle import surrogate as so


def test_gradients_match_finite_differences():
    rng = np.random.default_rng(0)
    p = so.init_params(1)
    n_pose, lna = 5, 7
    F = np.abs(rng.normal(0, 1, (n_pose * lna, so.N_FEAT)))
    pose = np.repeat(np.arange(n_pose), lna)
    target = rng.normal(0, 1, n_pose)
    loss, g = so.loss_and_grads(p, F, pose, target)
    eps = 1e-6
    for name in ("W1", "b1", "W2", "b2", "w3", "b3"):
        flat = p[name].reshape(-1)
        idx = rng.choice(flat.size, size=min(12, flat.size), replace=False)
        for i in idx:
            old = flat[i]
            flat[i] = old + eps
            lp, _ = so.loss_and_grads(p, F, pose, target)
            flat[i] = old - eps
            lm, _ = so.loss_and_grads(p, F, pose, target)
            flat[i] = old
            fd = (lp - lm) / (2 * eps)
            assert abs(fd - g[name].reshape(-1)[i]) <= 1e-6 * max(1.0, abs(fd)), (name, i)


def test_radial_features_continuous_at_cutoff_and_counts():
    prt = np.array([[0.0, 0.0, 0.0], [3.0, 0.0, 0.0]])
    t = np.array([1, 6])  # classes 1 and 2
    f_in = so.radial_features([[0, 0, 7.999999]], prt[:1], t[:1])
    f_out = so.radial_features([[0, 0, 8.000001]], prt[:1], t[:1])
    assert np.all(np.abs(f_in) < 1e-10) and np.all(f_out == 0)
    f = so.radial_features([[0.0, 0.0, 2.0]], prt, t)[0]
    d0, d1 = 2.0, np.hypot(3.0, 2.0)
    want1 = np.exp(-0.5 * (d0 - np.arange(1, 9)) ** 2) * 0.5 * (1 + np.cos(np.pi * d0 / 8))
    want2 = np.exp(-0.5 * (d1 - np.arange(1, 9)) ** 2) * 0.5 * (1 + np.cos(np.pi * d1 / 8))
    np.testing.assert_allclose(f[8:16], want1, rtol=1e-12)
    np.testing.assert_allclose(f[16:24], want2, rtol=1e-12)
    assert np.all(f[:8] == 0) and np.all(f[24:] == 0)


def test_interpolation_exact_on_nodes_and_linear_fields():
    rng = np.random.default_rng(1)
    n, h, o = 6, 0.7, np.array([1.0, -2.0, 0.5])
    grid = rng.normal(0, 1, (n, n, n, 3))
    for _ in range(10):
        i = rng.integers(0, n, 3)
        pos = o + h * i
        np.testing.assert_allclose(so.interp(grid, o, h, pos[None])[0], grid[i[0], i[1], i[2]],
                                   atol=1e-12)
    a = np.array([0.3, -1.2, 2.0])
    ax = o[:, None] + h * np.arange(n)[None, :]
    lin = (a[0] * ax[0][:, None, None] + a[1] * ax[1][None, :, None] + a[2] * ax[2][None, None, :])
    pos = o + rng.uniform(0, h * (n - 1), (20, 3))
    got = so.interp(lin[..., None], o, h, pos)[:, 0]
    np.testing.assert_allclose(got, pos @ a, rtol=1e-12, atol=1e-12)


def test_feature_layout():
    f = so.features([3, 0, 23], np.ones((3, 32)))
    assert f.shape == (3, 56)
    assert f[0, 3] == 1 and f[1, 0] == 1 and f[2, 23] == 1 and f[:, :24].sum() == 3
    assert np.all(f[:, 24:] == 1)
