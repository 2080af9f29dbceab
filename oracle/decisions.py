"""Decision-for-decision comparison of an engine run with the FP32 CPU path.

TEST INFRASTRUCTURE ONLY: used by tests/ and the smoke check.

The engine and the CPU path draw the same uniforms (Philox stream keyed by
seed, complex, replica and step) and decide with the same exp (decision_expf /
gdo_expf). Their FP32 energies differ only by rounding: the device's
summation order and hardware transcendentals against the CPU's libm. Two runs
therefore make identical accept and exchange decisions unless a decision's
margin, (E_new - E_old) * minus_beta - ln u for Metropolis
(src/kernel_cuda_l2_accept.cu:23) or (mb1 - mb2)(E2 - E1) - ln U for an
exchange (src/kernel_cuda_l1_exchangereplicas.cu:58-60), lies within that
rounding of zero.

`compare` walks every replica's accepted-step records (step and temperature
index of each) and finds its first divergent step. The divergence is
explained only if, in the CPU path's margin log, one of these holds:
  * the replica's own decision at that step (Metropolis; for delayed
    acceptance the composite test or a sub-step test of that group) has
    |margin| <= tol;
  * an exchange of the replica's ladder at or before that step had
    |margin| <= tol_x;
  * another replica of the same ladder diverged (explained) at or before that
    step: exchanges then compare a different energy.
tol = 4 |minus_beta|_max dE and tol_x = 4 max|mb_i - mb_j| dE, where dE is the
largest GPU-vs-CPU total-energy difference measured over the records both
runs share (floored at 2^-20 of the largest |E|). Anything else raises.

The accepted poses before a replica's divergence must also agree: bit for
bit with uniform moves (the reference's kind, move_scale * (2U - 1), one
product and one sum per coordinate on both sides), within 1e-4 with Gaussian
moves (the device's log, sqrt and sincos against libm's).
"""
import numpy as np


def _by_replica(recs, n_rep):
    out = [[] for _ in range(n_rep)]
    if len(recs) == 0:
        return [np.zeros(0, recs.dtype) for _ in range(n_rep)]
    order = np.lexsort((recs["step"], recs["replica"][:, 0]))
    r = recs[order]
    rid = r["replica"][:, 0]
    bounds = np.searchsorted(rid, np.arange(n_rep + 1))
    return [r[bounds[i]:bounds[i + 1]] for i in range(n_rep)]


def compare(cx, gpu_recs, cpu, nc, k_sub=1, gpu_state=None, da=False):
    """Check that every replica's decision sequence equals the CPU path's or
    diverges only where rounding explains it. `cpu` is an oracle McState run
    with enable_log and records; `gpu_recs` the engine's records of this
    complex. Returns a summary dict; raises AssertionError on an unexplained
    divergence or on a counter mismatch of an undisturbed replica or ladder."""
    n_rep = cx.n_rep
    n_lig, n_tmp = cx.n_lig, cx.n_tmp
    g_by = _by_replica(gpu_recs, n_rep)
    mb = np.asarray(cx.temp_minus_beta, np.float64)
    mb_max = float(np.abs(mb).max())
    dmb_max = float(np.abs(mb[:, None] - mb[None, :]).max())
    assert cpu.mlog is not None, "run the CPU path with enable_log"

    first = {}
    d_e, e_max = 0.0, 1e-30
    n_pose, pose_differ, pose_dmax = 0, 0, 0.0
    for r in range(n_rep):
        g, c = g_by[r], cpu.records(r)
        kg = np.stack([g["step"], g["replica"][:, 2]], 1) if len(g) else np.zeros((0, 2), int)
        kc = np.stack([c["step"], c["replica"][:, 2]], 1) if len(c) else np.zeros((0, 2), int)
        m = min(len(kg), len(kc))
        neq = np.nonzero(np.any(kg[:m] != kc[:m], axis=1))[0]
        k = int(neq[0]) if len(neq) else m
        if k > 0:
            eg = g["energy"]["e"][:k, 9].astype(np.float64)
            ec = c["energy"]["e"][:k, 9].astype(np.float64)
            d_e = max(d_e, float(np.abs(eg - ec).max()))
            e_max = max(e_max, float(np.abs(ec).max()))
            # the accepted poses up to the divergence: the same uniforms and
            # the same sums of offsets
            gm = np.ascontiguousarray(g["movematrix"][:k], np.float32)
            cm = np.ascontiguousarray(c["movematrix"][:k], np.float32)
            n_pose += k
            pose_differ += int(np.any(gm.view(np.uint32) != cm.view(np.uint32), axis=1).sum())
            pose_dmax = max(pose_dmax, float(np.abs(gm.astype(np.float64) - cm).max()))
        if k == len(kg) == len(kc):
            continue
        s_g = int(kg[k, 0]) if k < len(kg) else np.iinfo(np.int64).max
        s_c = int(kc[k, 0]) if k < len(kc) else np.iinfo(np.int64).max
        first[r] = min(s_g, s_c)
    d_e = max(d_e, e_max * 2.0 ** -20)
    tol = 4.0 * mb_max * d_e
    tol_x = 4.0 * dmb_max * d_e

    def ladder(r):
        p, t_l = divmod(r, n_tmp * n_lig)
        return p * n_lig + t_l % n_lig

    explained, why = {}, {}
    for r in sorted(first, key=lambda q: first[q]):
        s = first[r]
        i = s - cpu.log_step0
        reason = None
        if 0 <= i < cpu.log_steps:
            if not da and abs(cpu.mlog[r, i]) <= tol:
                reason = ("metropolis", float(cpu.mlog[r, i]))
            if da:
                # records sit at group ends: the group is steps s-k+1 .. s
                sub = cpu.mlog[r, max(0, i - k_sub + 1):i + 1]
                com = cpu.clog[r, i]
                if np.isfinite(com) and abs(com) <= tol:
                    reason = ("composite", float(com))
                elif np.any(np.abs(sub[np.isfinite(sub)]) <= tol):
                    reason = ("sub-step", float(np.nanmin(np.abs(sub))))
        lg = ladder(r)
        if reason is None and nc > 0:
            n_seg = min(cpu.log_segs, s // nc)
            if n_seg > 0:
                xm = np.abs(cpu.xlog[:n_seg, lg, :])
                hit = np.nonzero(np.isfinite(xm) & (xm <= tol_x))
                if len(hit[0]):
                    reason = ("exchange", float(cpu.xlog[hit[0][0], lg, hit[1][0]]))
        if reason is None:
            mates = [q for q in explained if ladder(q) == lg and explained[q] <= s]
            if mates:
                reason = ("ladder", mates[0])
        assert reason is not None, (
            f"replica {r}: decisions diverge at step {s} with no decision of it or its "
            f"ladder within rounding (tol {tol:.3g}, tol_x {tol_x:.3g}, dE {d_e:.3g}); "
            f"margin there {cpu.mlog[r, i] if 0 <= i < cpu.log_steps else None}")
        explained[r] = s
        why[r] = reason

    # poses: uniform moves (the reference's) are one product and one sum per
    # coordinate on both sides, so they match bit for bit; Gaussian moves go
    # through the device's and libm's log, sqrt and sincos and match to a few
    # ulps
    if getattr(cx, "move_kind", None) == 0:
        assert pose_differ == 0, (
            f"{pose_differ} of {n_pose} accepted poses differ from the CPU path's "
            f"(largest difference {pose_dmax:.3g})")
    assert pose_dmax <= 1e-4, pose_dmax

    # counters of undisturbed replicas and ladders
    if gpu_state is not None:
        dis_ladders = {ladder(r) for r in first}
        for r in range(n_rep):
            if r in first:
                continue
            assert gpu_state["acs_mc"][r] == cpu.acs_mc[r], (r, gpu_state["acs_mc"][r],
                                                              cpu.acs_mc[r])
            if ladder(r) not in dis_ladders:
                assert gpu_state["replica"][r, 2] == cpu.idx_tmp[r], r
                assert gpu_state["acs_x"][r] == cpu.acs_x[r], r
    return {"replicas": n_rep, "identical": n_rep - len(first), "explained": why,
            "d_e": d_e, "tol": tol, "tol_x": tol_x, "poses": n_pose,
            "poses_differing": pose_differ, "pose_max_diff": pose_dmax}
