"""The engine on complexes read from the reference's own input files
(tests/golden/inputs): 1a07C1 (21 ligand conformers, 3 receptor models) and
10gsA00 (51 ligand conformers, 4,408 KDE points, 196 MCS templates).

  - every pose of data/trajectory.out (a reference run on 1a07C1) evaluated
    by gd_calc_energy reproduces the MCC and total energy the reference
    printed;
  - every energy term against the FP64 oracle (same bound as the synthetic
    parity tests);
  - an MC run follows the FP32 CPU path decision for decision.
"""
import os

import numpy as np
import pytest

from geaxdock_gpu_amd import io

pytestmark = pytest.mark.gpu

G = os.path.join(os.path.dirname(__file__), "golden", "inputs")
PARAS = os.path.join(G, "parameters", "paras")
A07 = [os.path.join(G, "1a07C1", f) for f in ("1a07C1.sdf", "1a07C.pdb", "1a07C1-0.8.ff")]
GS10 = [os.path.join(G, "10gs", f) for f in ("10gs_ligand.sdf", "10gsA00.pdb", "10gsA00.ff")]


@pytest.fixture(scope="module")
def gd():
    import geaxdock_gpu_amd as g
    return g


def trajectory():
    with open(os.path.join(G, "trajectory.out")) as f:
        rows = [l.split() for l in f.read().split("\n")[1:] if l.strip()]
    ip = np.array([int(r[0]) for r in rows], np.int32)
    il = np.array([int(r[1]) for r in rows], np.int32)
    mm = np.array([[float(v) for v in r[2:8]] for r in rows], np.float32)
    cms = np.array([float(r[8]) for r in rows])
    ener = np.array([float(r[9]) for r in rows])
    return ip, il, mm, cms, ener


@pytest.mark.parametrize("split", [1, 2])
def test_reference_trajectory_on_gpu(gd, split):
    """2,666 poses of data/trajectory.out: total energy within 2e-5 of the
    printed value (printed to 1e-6) and MCC within 2e-6."""
    cx = io.load_complex(*A07, PARAS, complex_id="1a07C1")
    ip, il, mm, cms, ener = trajectory()
    with gd.Engine(seed=1, receptor_split=split) as eng:
        cid = eng.add_complex(cx)
        out = eng.calc_energy(cid, il, ip, mm)
    de = np.abs(out["e"][:, 9].astype(np.float64) - ener)
    dc = np.abs(out["cms"].astype(np.float64) - cms)
    print(f"\nsplit {split}: worst |d ener| {de.max():.2e} at {de.argmax()}, "
          f"worst |d cms| {dc.max():.2e}")
    assert de.max() <= 2e-5, (de.argmax(), out["e"][de.argmax(), 9], ener[de.argmax()])
    assert dc.max() <= 2e-6, (dc.argmax(), out["cms"][dc.argmax()], cms[dc.argmax()])


def _poses(rng, n, t_sigma):
    mm = np.zeros((n, 6), np.float32)
    mm[:, :3] = rng.normal(0.0, t_sigma, size=(n, 3))
    mm[:, 3:] = rng.uniform(-np.pi, np.pi, size=(n, 3))
    mm[0] = 0.0
    return mm


@pytest.mark.parametrize("which,split", [("1a07C1", 0), ("1a07C1", 8), ("10gsA00", 0),
                                         ("10gsA00", 1), ("10gsA00", 2), ("10gsA00", 4),
                                         ("10gsA00", 8)])
def test_real_complex_energy_terms_vs_fp64(oracle, gd, which, split):
    """Every term against FP64 on the reference's complexes. The receptor
    split's groups share out the receptor points and also each atom's KDE
    points (4,408 in 10gsA00), reference contacts and the MCS templates (196
    in 10gsA00), so the KDE, MCS and MCC terms are checked at every split."""
    from test_gpu_parity import check_energy_parity
    files, cid_name = (A07, "1a07C1") if which == "1a07C1" else (GS10, "10gsA00")
    cx = io.load_complex(*files, PARAS, complex_id=cid_name)
    rng = np.random.default_rng(17)
    n = 40
    mm = _poses(rng, n, 1.0)
    il = rng.integers(0, cx.n_lig, size=n).astype(np.int32)
    ip = rng.integers(0, cx.n_prt, size=n).astype(np.int32)
    il[0] = ip[0] = 0
    with gd.Engine(seed=2, receptor_split=split) as eng:
        cid = eng.add_complex(cx)
        rep = {}
        check_energy_parity(oracle, eng, cid, cx, mm, il, ip, rep)
    print(f"\n{which} split {split} worst (rel, cond):",
          {k: (f"{v[0]:.2e}", f"{v[1]:.2e}") for k, v in rep.items()})


@pytest.mark.parametrize("which,n_tmp", [("1a07C1", 4), ("10gsA00", 2)])
def test_real_complex_mc_tracks_cpu_path(oracle, gd, which, n_tmp):
    """The reference driver's move scales (0.02, 0.08) and a temperature
    ladder on a real complex: GPU and FP32 CPU path make the same decisions."""
    from geaxdock_gpu_amd import abi
    files, cid_name = (A07, "1a07C1") if which == "1a07C1" else (GS10, "10gsA00")
    cx = io.load_complex(*files, PARAS, complex_id=cid_name, n_tmp=n_tmp, floor_temp=0.044,
                         ceiling_temp=0.36)
    n_steps, nc = 150, 10
    recs = []
    with gd.Engine(seed=11) as eng:
        cid = eng.add_complex(cx)
        eng.run(n_steps, nc, sink=lambda c, r: recs.append(r))
        state = eng.state(cid)
    g_all = np.concatenate(recs) if recs else np.zeros(0, abi.RECORD_DTYPE)
    d = cx.desc()
    cpu = oracle.McState(cx, rec_cap=n_steps + 1)
    cpu.enable_log(0, n_steps, n_steps // nc)
    oracle.mc_run(d, cpu, seed=11, init=True, step0=0, n_steps=0, nc=nc)
    oracle.mc_run(d, cpu, seed=11, init=False, step0=0, n_steps=n_steps, nc=nc)
    from oracle import decisions
    out = decisions.compare(cx, g_all, cpu, nc, gpu_state=state)
    print(f"\n{which}: {out['identical']}/{cx.n_rep} replicas with identical decisions, "
          f"explained {out['explained']}; acceptance "
          f"{state['acs_mc'].sum() / (cx.n_rep * n_steps):.3f}")
    assert state["acs_mc"].sum() > 0
