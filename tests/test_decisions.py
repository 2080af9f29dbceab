"""The decision-parity checker (oracle/decisions.py) on CPU-only runs.

Two FP32 CPU paths on receptors whose coordinates differ by N(0, 3e-3 A) stand
in for the engine: their energies differ by a few 1e-5, so every divergence
must be explained by a near-threshold decision. A run with a planted decision bug (one
temperature decided with its neighbour's minus_beta, as a wrong beta after an
exchange would) must be rejected, and so must accepted poses one ulp off the
CPU path's under uniform moves."""
import copy

import numpy as np
import pytest

from geaxdock_gpu_amd import synth
from oracle import decisions, gdo


def _run(cx, seed, n_steps, nc, log=False):
    d = cx.desc()
    st = gdo.McState(cx, rec_cap=n_steps + 1)
    if log:
        st.enable_log(0, n_steps, n_steps // nc if nc else 0)
    gdo.mc_run(d, st, seed=seed, init=True, step0=0, n_steps=0, nc=nc)
    gdo.mc_run(d, st, seed=seed, init=False, step0=0, n_steps=n_steps, nc=nc)
    recs = np.concatenate([st.records(r) for r in range(cx.n_rep)])
    return st, recs


def _state(st):
    rep = np.zeros((len(st.idx_tmp), 4), np.int32)
    rep[:, 2] = st.idx_tmp
    return {"acs_mc": st.acs_mc, "acs_x": st.acs_x, "replica": rep}


def test_identical_runs_are_identical():
    cx = synth.make_complex(3, lna=20, pnp=600, n_lig=2, n_prt=1, n_tmp=6)
    cpu, _ = _run(cx, 5, 200, 20, log=True)
    other, recs = _run(cx, 5, 200, 20)
    out = decisions.compare(cx, recs, cpu, 20, gpu_state=_state(other))
    assert out["identical"] == cx.n_rep and not out["explained"]


def test_rounding_level_differences_are_explained():
    """Receptors 3e-3 A apart, cold ladder (kT 0.01-0.1): over three seeds a
    replica diverges, and each divergence is explained."""
    cx = synth.make_complex(4, lna=24, pnp=700, n_lig=8, n_prt=1, n_tmp=8, t_floor=0.01,
                            t_ceiling=0.1)
    cy = copy.copy(cx)
    cy.prt_x = (cx.prt_x + np.random.default_rng(0).normal(0, 3e-3, cx.prt_x.shape)).astype(
        np.float32)
    diverged = 0
    for seed in range(3):
        cpu, _ = _run(cx, seed, 400, 20, log=True)
        other, recs = _run(cy, seed, 400, 20)
        out = decisions.compare(cx, recs, cpu, 20, gpu_state=_state(other))
        diverged += len(out["explained"])
    print("\nexplained divergences over 3 seeds:", diverged)
    assert diverged > 0


def test_planted_decision_bug_is_rejected():
    cx = synth.make_complex(4, lna=24, pnp=700, n_lig=2, n_prt=1, n_tmp=8)
    bad = copy.copy(cx)
    bad.temp_minus_beta = cx.temp_minus_beta.copy()
    bad.temp_minus_beta[2] = cx.temp_minus_beta[3]
    cpu, _ = _run(cx, 1, 300, 20, log=True)
    _, recs = _run(bad, 1, 300, 20)
    with pytest.raises(AssertionError, match="diverge"):
        decisions.compare(cx, recs, cpu, 20)


def test_planted_pose_error_is_rejected():
    """Uniform moves must give the CPU path's poses bit for bit: a run whose
    accepted poses are off by one ulp in one coordinate (decisions unchanged)
    is rejected; with Gaussian moves such a difference is within bounds, and
    one of 1e-3 is not."""
    from geaxdock_gpu_amd import abi
    for kind in (abi.MOVE_UNIFORM, abi.MOVE_GAUSSIAN):
        cx = synth.make_complex(6, lna=20, pnp=600, n_lig=2, n_prt=1, n_tmp=6, move_kind=kind)
        cpu, _ = _run(cx, 3, 200, 20, log=True)
        other, recs = _run(cx, 3, 200, 20)
        out = decisions.compare(cx, recs, cpu, 20, gpu_state=_state(other))
        assert out["poses"] > 0 and out["poses_differing"] == 0
        bad = recs.copy()
        bad["movematrix"][:, 0] = np.nextafter(bad["movematrix"][:, 0], np.float32(np.inf))
        if kind == abi.MOVE_UNIFORM:
            with pytest.raises(AssertionError, match="accepted poses differ"):
                decisions.compare(cx, bad, cpu, 20, gpu_state=_state(other))
        else:
            out = decisions.compare(cx, bad, cpu, 20, gpu_state=_state(other))
            assert out["poses_differing"] > 0
            bad["movematrix"][:, 1] += np.float32(1e-3)
            with pytest.raises(AssertionError):
                decisions.compare(cx, bad, cpu, 20, gpu_state=_state(other))
