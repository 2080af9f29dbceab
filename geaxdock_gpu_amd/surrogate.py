"""Learned surrogate of the receptor-pair energy and its delayed-acceptance use.

The network is a per-ligand-atom MLP 56 -> 128 -> 128 -> 1 (SiLU), summed over
the atoms of a pose. Its 56 inputs are a one-hot of the atom's type (24) and
32 radial receptor-density features (4 receptor classes x 8 Gaussian shells,
read from per-conformer grids by trilinear interpolation). It learns the
weighted sum of the six receptor-pair terms (vdW, ele, PMF, PSP, HDB, HPC) in
standardized units; the delayed-acceptance kernel adds the exact KDE, MCS
and pocket-distance terms back, so the surrogate total is
    S = scale * sum_atoms MLP(features) + cheap terms.

Parameter vector layout (gd_sur_set_params): W1 as a [64][128] matrix whose
rows 0..55 are the input weights, row 56 the first-layer bias and rows 57..63
zero; then W2 [128][128], b2 [128], w3 [128], b3 [1].
"""
import numpy as np

N_FEAT = 56
N_IN = 64
N_HID = 128
N_PARAM = N_IN * N_HID + N_HID * N_HID + N_HID + N_HID + 1

PAIR_TERMS = slice(0, 6)
CHEAP_TERMS = slice(6, 9)


def init_params(seed=0):
    """He-scaled normal weights, zero biases (dict of float64 arrays)."""
    rng = np.random.default_rng(seed)
    return {"W1": rng.normal(0, np.sqrt(2.0 / N_FEAT), (N_FEAT, N_HID)),
            "b1": np.zeros(N_HID),
            "W2": rng.normal(0, np.sqrt(2.0 / N_HID), (N_HID, N_HID)),
            "b2": np.zeros(N_HID),
            "w3": rng.normal(0, np.sqrt(1.0 / N_HID), N_HID),
            "b3": np.zeros(1)}


def pack_params(p):
    """dict -> flat float32 vector in the engine's layout."""
    v = np.zeros(N_PARAM, np.float32)
    W1 = np.zeros((N_IN, N_HID))
    W1[:N_FEAT] = p["W1"]
    W1[N_FEAT] = p["b1"]
    o = 0
    v[o:o + N_IN * N_HID] = W1.reshape(-1)
    o += N_IN * N_HID
    v[o:o + N_HID * N_HID] = np.asarray(p["W2"]).reshape(-1)
    o += N_HID * N_HID
    v[o:o + N_HID] = p["b2"]
    o += N_HID
    v[o:o + N_HID] = p["w3"]
    o += N_HID
    v[o] = np.asarray(p["b3"]).reshape(-1)[0]
    return v


def unpack_params(v):
    """flat vector (parameters or gradients) -> dict."""
    v = np.asarray(v)
    o = N_IN * N_HID
    W1 = v[:o].reshape(N_IN, N_HID)
    out = {"W1": W1[:N_FEAT], "b1": W1[N_FEAT]}
    out["W2"] = v[o:o + N_HID * N_HID].reshape(N_HID, N_HID)
    o += N_HID * N_HID
    out["b2"] = v[o:o + N_HID]
    o += N_HID
    out["w3"] = v[o:o + N_HID]
    o += N_HID
    out["b3"] = v[o:o + 1]
    return out


def pair_energy(energy, enepara):
    """Weighted receptor-pair part of the total: sum_{i<6} w_i e_i."""
    w = np.asarray(enepara.w, np.float64)[PAIR_TERMS]
    return np.asarray(energy["e"], np.float64)[:, PAIR_TERMS] @ w


def sample_poses(eng, cid, n, rng, t_sigma=2.0, from_state=True, jitter=(0.5, 0.3)):
    """Training poses: the replicas' current poses jittered by a normal
    offset (translation, rotation), plus uniformly random orientations around
    the pocket centre. Returns (idx_lig, idx_prt, movematrix)."""
    cx = eng.complexes[cid]
    n_state = n // 2 if from_state else 0
    mm = np.zeros((n, 6), np.float32)
    il = rng.integers(0, cx.n_lig, n).astype(np.int32)
    ip = rng.integers(0, cx.n_prt, n).astype(np.int32)
    if n_state:
        st = eng.state(cid)
        pick = rng.integers(0, len(st["movematrix"]), n_state)
        mm[:n_state] = st["movematrix"][pick]
        mm[:n_state, :3] += rng.normal(0, jitter[0], (n_state, 3))
        mm[:n_state, 3:] += rng.normal(0, jitter[1], (n_state, 3))
        il[:n_state] = st["replica"][pick, 3]
        ip[:n_state] = st["replica"][pick, 1]
    r = n - n_state
    mm[n_state:, :3] = rng.normal(0, t_sigma, (r, 3))
    mm[n_state:, 3:] = rng.uniform(-np.pi, np.pi, (r, 3))
    return il, ip, mm


def build_dataset(eng, cid, n, seed=0, **kw):
    """Label n poses with the exact engine energy and append them to the
    engine's training set. Returns (mean, std) of the pair-energy target,
    which is stored standardized."""
    rng = np.random.default_rng(seed)
    il, ip, mm = sample_poses(eng, cid, n, rng, **kw)
    e = eng.calc_energy(cid, il, ip, mm)
    y = pair_energy(e, eng.complexes[cid].enepara_struct)
    ok = np.isfinite(y)
    # clip the steep repulsive wall: the Metropolis test rejects such poses
    # whatever the exact value, so they only need to rank high
    cap = np.percentile(y[ok], 95) if ok.any() else 0.0
    y = np.where(ok, np.minimum(y, cap), cap)
    mu, sd = float(y.mean()), float(y.std() + 1e-12)
    eng.sur_dataset_add(cid, il, ip, mm, ((y - mu) / sd).astype(np.float32))
    return mu, sd


def fit(eng, n_train, n_test, epochs=40, batch_poses=256, lr=2e-3, seed=0):
    """Train on rows [0, n_train), report RMSE (standardized units) on the
    held-out rows [n_train, n_train + n_test)."""
    eng.sur_set_params(pack_params(init_params(seed)))
    hist, secs = eng.sur_train(0, n_train, epochs, batch_poses, lr)
    lte, _ = eng.sur_loss_grad(n_train, n_train + n_test, grad=False)
    return dict(train_loss=float(hist[-1]), test_rmse=float(np.sqrt(lte)), seconds=secs,
                history=hist)
