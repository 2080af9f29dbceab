"""FP64 oracle of the learned energy surrogate. TEST INFRASTRUCTURE ONLY:
imported by tests/ and the benchmark's checks, never by the engine.

Model (the engine's definition, restated in float64 with numpy):

  per ligand atom l at position r_l, 56 features
    f[0:24]          one-hot of the ligand atom type
    f[24 + 8c + k]   sum over receptor points p of class c = t_p mod 4 with
                     d = |r_l - r_p| < 8 A of exp(-(d - (k+1))^2 / 2) * fc(d),
                     fc(d) = (1 + cos(pi d / 8)) / 2,   k = 0..7
  the receptor part of the 32 radial features is read from a grid (spacing h,
  origin o, n^3 points per protein conformer) by trilinear interpolation,
  positions clamped to the grid;
  MLP  h1 = silu(f W1 + b1)      W1: 56 x 128
       h2 = silu(h1 W2 + b2)     W2: 128 x 128
       y  = h2 . w3 + b3         w3: 128
  pose value S = sum_l y_l; training loss = mean over poses (S - target)^2.
"""
import numpy as np

N_TYPES = 24
N_CLASS = 4
N_RBF = 8
N_FEAT = N_TYPES + N_CLASS * N_RBF  # 56
HIDDEN = 128
CUTOFF = 8.0


def silu(x):
    return x / (1.0 + np.exp(-x))


def dsilu(x):
    s = 1.0 / (1.0 + np.exp(-x))
    return s * (1.0 + x * (1.0 - s))


def radial_features(pos, prt_xyz, prt_t):
    """Direct (no grid) 32 radial features at positions pos [n,3] (float64)."""
    pos = np.asarray(pos, np.float64).reshape(-1, 3)
    P = np.asarray(prt_xyz, np.float64).reshape(-1, 3)
    cls = np.asarray(prt_t) % N_CLASS
    out = np.zeros((len(pos), N_CLASS * N_RBF))
    centers = np.arange(1, N_RBF + 1, dtype=np.float64)
    for i, r in enumerate(pos):
        d = np.linalg.norm(P - r, axis=1)
        m = d < CUTOFF
        dm, cm = d[m], cls[m]
        fc = 0.5 * (1.0 + np.cos(np.pi * dm / CUTOFF))
        g = np.exp(-0.5 * (dm[:, None] - centers[None, :]) ** 2) * fc[:, None]
        for c in range(N_CLASS):
            out[i, c * N_RBF:(c + 1) * N_RBF] = g[cm == c].sum(axis=0)
    return out


def grid_features(origin, h, n, prt_xyz, prt_t):
    """Radial features on an n^3 grid, [n, n, n, 32] indexed [ix, iy, iz]."""
    ax = origin[0] + h * np.arange(n)
    ay = origin[1] + h * np.arange(n)
    az = origin[2] + h * np.arange(n)
    X, Y, Z = np.meshgrid(ax, ay, az, indexing="ij")
    pts = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)
    return radial_features(pts, prt_xyz, prt_t).reshape(n, n, n, -1)


def interp(grid, origin, h, pos):
    """Trilinear interpolation of grid [n,n,n,F] at pos [m,3], clamped."""
    n = grid.shape[0]
    u = (np.asarray(pos, np.float64) - np.asarray(origin, np.float64)) / h
    u = np.clip(u, 0.0, n - 1 - 1e-9)
    i0 = np.floor(u).astype(int)
    t = u - i0
    out = np.zeros((len(u), grid.shape[-1]))
    for dx in (0, 1):
        for dy in (0, 1):
            for dz in (0, 1):
                w = ((t[:, 0] if dx else 1 - t[:, 0]) * (t[:, 1] if dy else 1 - t[:, 1]) *
                     (t[:, 2] if dz else 1 - t[:, 2]))
                out += w[:, None] * grid[i0[:, 0] + dx, i0[:, 1] + dy, i0[:, 2] + dz]
    return out


def features(lig_t, radial):
    f = np.zeros((len(lig_t), N_FEAT))
    f[np.arange(len(lig_t)), np.asarray(lig_t)] = 1.0
    f[:, N_TYPES:] = radial
    return f


def init_params(seed=0):
    rng = np.random.default_rng(seed)
    return {
        "W1": rng.normal(0, np.sqrt(2.0 / N_FEAT), (N_FEAT, HIDDEN)),
        "b1": np.zeros(HIDDEN),
        "W2": rng.normal(0, np.sqrt(2.0 / HIDDEN), (HIDDEN, HIDDEN)),
        "b2": np.zeros(HIDDEN),
        "w3": rng.normal(0, np.sqrt(1.0 / HIDDEN), HIDDEN),
        "b3": np.zeros(1),
    }


def forward(params, F, pose_of_atom, n_pose):
    """F [N,56] atom features, pose_of_atom [N] -> pose values [n_pose] and cache."""
    H1 = F @ params["W1"] + params["b1"]
    A1 = silu(H1)
    H2 = A1 @ params["W2"] + params["b2"]
    A2 = silu(H2)
    y = A2 @ params["w3"] + params["b3"][0]
    S = np.zeros(n_pose)
    np.add.at(S, pose_of_atom, y)
    return S, (F, H1, A1, H2, A2)


def loss_and_grads(params, F, pose_of_atom, target):
    """Mean squared error over poses and its gradients (FP64)."""
    n_pose = len(target)
    S, (F, H1, A1, H2, A2) = forward(params, F, pose_of_atom, n_pose)
    r = S - target
    loss = np.mean(r * r)
    dS = 2.0 * r / n_pose
    dy = dS[pose_of_atom]
    g = {"w3": A2.T @ dy, "b3": np.array([dy.sum()])}
    dA2 = np.outer(dy, params["w3"])
    dH2 = dA2 * dsilu(H2)
    g["W2"] = A1.T @ dH2
    g["b2"] = dH2.sum(0)
    dA1 = dH2 @ params["W2"].T
    dH1 = dA1 * dsilu(H1)
    g["W1"] = F.T @ dH1
    g["b1"] = dH1.sum(0)
    return loss, g


def rows(params, F):
    """Per-row network outputs y for feature rows F [N, 56]."""
    S, _ = forward(params, F, np.arange(len(F)), len(F))
    return S


def tabulate(params, grid, types):
    """Network output at every grid node (features = the node's values) for
    each ligand atom type: {type: [n, n, n]}."""
    n = grid.shape[0]
    radial = grid.reshape(-1, grid.shape[-1])
    out = {}
    for t in types:
        F = features(np.full(len(radial), int(t)), radial)
        out[int(t)] = rows(params, F).reshape(n, n, n)
    return out
