"""Learned surrogate of the receptor-pair energy and its delayed-acceptance use.

The network is a per-ligand-atom MLP 56 -> 128 -> 128 -> 1 (SiLU), summed over
the atoms of a pose. Its 56 inputs are a one-hot of the atom's type (24) and
32 radial receptor-density features (4 receptor classes x 8 Gaussian shells,
read from per-conformer grids by trilinear interpolation). It learns the
weighted sum of the six receptor-pair terms (vdW, ele, PMF, PSP, HDB, HPC) in
standardized units; the delayed-acceptance kernel adds the exact KDE, MCS
and pocket-distance terms back, so the surrogate total is
    S = scale * sum_atoms MLP(features) / lna + cheap terms.

Parameter vector layout (gd_sur_set_params): W1 as a [64][128] matrix whose
rows 0..55 are the input weights, row 56 the first-layer bias and rows 57..63
zero; then W2 [128][128], b2 [128], w3 [128], b3 [1].
"""
import numpy as np

from . import abi

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


def pair_target(energy, enepara, lna):
    """Receptor-pair part of a pose's total as a sum over its ligand atoms:
    T = lna * sum_{i<6} w_i (e_i - b_i) = sum_{i<6} w_i a_i raw_i, where the
    normalized terms are e_i = a_i raw_i / lna + b_i."""
    w = np.asarray(enepara.w, np.float64)[PAIR_TERMS]
    b = np.asarray(enepara.b_para, np.float64)[PAIR_TERMS]
    e = np.asarray(energy["e"], np.float64)[:, PAIR_TERMS]
    return np.asarray(lna, np.float64) * ((e - b) @ w)


def perturb(mm, move_scale, move_kind, n_moves, rng):
    """Apply n_moves random MC moves of the engine's proposal kind."""
    mm = np.array(mm, np.float32, copy=True)
    sc = np.asarray(move_scale, np.float32)
    for _ in range(int(n_moves)):
        if move_kind == 1:
            mm += (rng.normal(0.0, 1.0, mm.shape) * sc).astype(np.float32)
        else:
            mm += ((2.0 * rng.random(mm.shape) - 1.0) * sc).astype(np.float32)
    return mm


class PoseLog:
    """Bounded reservoir of accepted states from an engine's record sink."""

    def __init__(self, capacity, seed=0):
        self.cap = int(capacity)
        self.rng = np.random.default_rng(seed)
        self.seen = 0
        self.buf = None

    def __call__(self, cid, recs):
        n = len(recs)
        if n == 0:
            return
        take = recs
        if self.buf is None:
            self.buf = np.zeros(self.cap, recs.dtype)
            self.cid = np.zeros(self.cap, np.int32)
            self.fill = 0
        # reservoir sampling, vectorized per window
        idx = self.seen + np.arange(n)
        room = max(0, self.cap - self.fill)
        if room:
            k = min(room, n)
            self.buf[self.fill:self.fill + k] = take[:k]
            self.cid[self.fill:self.fill + k] = cid
            self.fill += k
            take, idx = take[k:], idx[k:]
        if len(take):
            j = (self.rng.random(len(take)) * (idx + 1)).astype(np.int64)
            hit = j < self.cap
            self.buf[j[hit]] = take[hit]
            self.cid[j[hit]] = cid
        self.seen += n

    def records(self):
        if self.buf is None:
            return None, None
        return self.buf[:self.fill], self.cid[:self.fill]


def log_states(eng, n_snapshots, every, nc=20, with_time=False):
    """Accepted states of a plain run: after every `every` steps, each
    replica's current pose and its exact energy (a record array with the
    gd_record layout, and the complex of each entry). Memory is bounded by
    n_snapshots x replicas, whatever the run length. With `with_time`, the
    run's summed kernel seconds are returned as a third value."""
    out, cids = [], []
    t_kernel = 0.0
    for _ in range(n_snapshots):
        t_kernel += eng.run(every, nc)["t_compute"]
        for c in range(len(eng.complexes)):
            st = eng.state(c)
            r = np.zeros(len(st["movematrix"]), abi.RECORD_DTYPE)
            r["replica"] = st["replica"]
            r["energy"] = st["energy"]
            r["movematrix"] = st["movematrix"]
            out.append(r)
            cids.append(np.full(len(r), c, np.int32))
    if with_time:
        return np.concatenate(out), np.concatenate(cids), t_kernel
    return np.concatenate(out), np.concatenate(cids)


def build_training_set(eng, recs, rec_cid, n_poses, perturbed=0.5, max_moves=4, seed=0):
    """Poses for training: logged accepted states, and the same states moved
    by 1..max_moves proposals (the surrogate ranks proposals). Labels are the
    exact engine energies. Returns per-pose arrays (shuffled)."""
    rng = np.random.default_rng(seed)
    pick = rng.integers(0, len(recs), n_poses)
    r = recs[pick]
    cid = rec_cid[pick].astype(np.int32)
    il = r["replica"][:, 3].astype(np.int32)
    ip = r["replica"][:, 1].astype(np.int32)
    mm = r["movematrix"].astype(np.float32)
    energy = r["energy"].copy()
    n_pert = int(n_poses * perturbed)
    for c in np.unique(cid[:n_pert]):
        sel = np.nonzero(cid[:n_pert] == c)[0]
        cx = eng.complexes[int(c)]
        moves = rng.integers(1, max_moves + 1, len(sel))
        for m in np.unique(moves):
            s2 = sel[moves == m]
            mm[s2] = perturb(mm[s2], cx.move_scale, cx.move_kind, m, rng)
        energy[sel] = eng.calc_energy(int(c), il[sel], ip[sel], mm[sel])
    lna = np.array([eng.complexes[int(c)].lna for c in cid], np.float64)
    T = np.zeros(n_poses)
    for c in np.unique(cid):
        sel = cid == c
        T[sel] = pair_target(energy[sel], eng.complexes[int(c)].enepara_struct, lna[sel])
    ok = np.isfinite(T)
    order = rng.permutation(np.nonzero(ok)[0])
    return dict(cid=cid[order], idx_lig=il[order], idx_prt=ip[order], movematrix=mm[order],
                target=T[order], lna=lna[order], energy=energy[order])


def load_training_set(eng, ts, clip_pct=99.5):
    """Standardize T per atom, t = (T - lna m) / sigma, clip the repulsive tail,
    and append the poses to the engine's dataset in order. Returns (m, sigma):
    the engine's surrogate scale is sigma."""
    T, lna = ts["target"], ts["lna"]
    m = float(np.mean(T / lna))
    t = T - lna * m
    sigma = float(np.std(t) + 1e-12)
    t = t / sigma
    cap = np.percentile(t, clip_pct)
    t = np.minimum(t, cap).astype(np.float32)
    ts["t"] = t
    eng.sur_dataset_reserve(int(eng.sur_dataset_size() + lna.sum() + 64))
    eng.sur_dataset_add_multi(ts["cid"], ts["idx_lig"], ts["idx_prt"], ts["movematrix"], t)
    return m, sigma


def fit(eng, ts, sigma, n_train, epochs=8, batch_poses=4096, lr=1e-3, seed=0):
    """Train on poses [0, n_train) of the engine's dataset and evaluate the
    held-out rest: RMSE in standardized units and of the pair energy per the
    engine's normalization (sum_i w_i e_i units)."""
    eng.sur_set_params(pack_params(init_params(seed)))
    n = len(ts["t"])
    hist, secs = eng.sur_train(0, n_train, epochs, batch_poses, lr)
    pred = eng.sur_predict(n_train, n)
    t = ts["t"][n_train:]
    lna = ts["lna"][n_train:]
    err = (pred.astype(np.float64) - t) * sigma / lna
    keep = t < t.max()  # clipped poses excluded from the energy RMSE
    return dict(train_loss=float(hist[-1]), history=[float(h) for h in hist], seconds=secs,
                test_rmse_std=float(np.sqrt(np.mean((pred - t) ** 2))),
                test_rmse_energy=float(np.sqrt(np.mean(err[keep] ** 2))),
                test_energy_std=float(np.std((ts["target"][n_train:] / lna)[keep])),
                n_train=int(n_train), n_test=int(n - n_train))
