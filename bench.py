#!/usr/bin/env python3
"""Benchmark: MC replica-steps/s, plain vs surrogate delayed acceptance, on one MI355X.

Workload (default, `--config batch --batch 1024`): 1,024 synthetic complexes,
ligand atoms U[12,64], receptor points U[800,4000], each 5 ligand x 3 protein
conformers x 8 temperatures (122,880 replicas), an exchange attempt every 20
MC steps, FP32, batched persistent graph-captured launches; delayed
acceptance with up to k = 4 surrogate-screened proposals per exact
evaluation, compared with the plain engine on the same complexes, seed and
start state. `--config ensemble` is the 3,520-replica single-complex
ensemble, `--config single` the CPU-sized complex.

Step unit: one bench step is `--mc-per-step` (default 250) MC steps of every
replica, so `--steps 20 --warmup 5` times 5,000 MC steps (the workload's run
length) after 1,250 MC steps of warm-up. MC step counts are rounded up to a
multiple of lcm(k, nc): delayed acceptance advances in whole groups of k
sub-steps and exchanges on group boundaries. The counts used are reported
(`mc_steps`, `mc_warmup`).

Default mode (da):
 1. a plain warm-up whose replica states are snapshotted for training;
 2. training poses (snapshots, and snapshots moved by 1-4 proposals) labelled
    by the exact energy; the MLP trained on the GPU; held-out RMSE;
 3. from one checkpoint, without records: the DA run (its replica-steps per
    second of wall time is `value`), the DA run with the network removed
    (scale 0), and the plain run. The plain run is taken in segments, so its
    per-complex best energies are read at the DA run's number of exact
    evaluations (steps / k), at the DA run's wall time and at equal steps.

roofline: the plain MC kernel, FP32 VALU bound. Algorithmic work is 36 flops
per ligand-atom x receptor-point pair per replica-step (distance 10, soft LJ
13, far-field electrostatics 5, PMF 8; conditional terms not counted), so
36 * sum(n_rep * lna * pnp) per MC step. `achieved` divides the plain run's
flops by its summed launch durations (HIP events on the engine's stream around
each window's launch, all per-APL kernel classes of the window included).
Peak: MI355X FP32 vector 157.3 TFLOP/s (78.6 without packed FP32).
`traffic` is the HBM bytes per launch of the plain kernel from the workload's
FETCH_SIZE/WRITE_SIZE counter profile (PLAIN_TRAFFIC: a per-launch part plus a
per-MC-step part), null for workloads and launch shapes not profiled.

cpu_baseline: the FP32 CPU restatement (oracle/, OpenMP over replicas) on a
bounded sample of the same workload, on this host's cores.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FLOPS_PER_PAIR = 36.0
PEAK_FP32 = 157.3e12
PEAK_FP32_UNPACKED = 78.6e12
PEAK_BF16 = 2.5e15
SUR_FWD_FLOPS_PER_ROW = 2.0 * (64 * 128 + 128 * 128 + 128)
SUR_TRAIN_FLOPS_PER_ROW = SUR_FWD_FLOPS_PER_ROW + 2.0 * (128 * 128 + 128 * 128 + 64 * 128)
METRIC = "MC replica-steps/s, plain vs surrogate-DA, at matched best-pose energy"
# HBM bytes of one plain MC launch: (read per launch, read per MC step,
# written per launch, written per MC step). From rocprofv3 FETCH_SIZE and
# WRITE_SIZE of eager 260- and 500-step launches of the same workload, each
# scaled by what the counter reads for a copy kernel of known size (2.000 KiB
# per FETCH_SIZE unit, 1.000 KiB per WRITE_SIZE unit). The per-launch part is
# every workgroup staging its receptor, tables and KDE points into LDS; the
# per-step part is L2 misses of MCS templates, contact lists and PSP rows.
PLAIN_TRAFFIC = {
    ("batch", 1024): (9.111e8, 4.164e5, 1.576e7, 238.5),
    ("ensemble", None): (8.04e6, 0.0, 8.01e5, 0.0),
}


def parse(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20, help="timed bench steps")
    ap.add_argument("--warmup", type=int, default=5, help="warm-up bench steps")
    ap.add_argument("--mc-per-step", type=int, default=250,
                    help="MC steps of every replica per bench step")
    ap.add_argument("--nc", type=int, default=20, help="MC steps per exchange")
    ap.add_argument("--config", default="batch", choices=["ensemble", "single", "batch"])
    ap.add_argument("--batch", type=int, default=1024, help="complexes in the batch workload")
    ap.add_argument("--cpu-seconds", type=float, default=12.0,
                    help="length of the CPU-path sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--lpr", type=int, default=0)
    ap.add_argument("--apl", type=int, default=0)
    ap.add_argument("--split", type=int, default=-1,
                    help="receptor split (0 = engine's per-workgroup choice, or 1, 2, 4, 8)")
    ap.add_argument("--ladders", type=int, default=-1,
                    help="temperature ladders per workgroup (0 = engine's choice)")
    ap.add_argument("--launch", type=int, default=-1, help="GD_LAUNCH_* flags")
    ap.add_argument("--tile", type=int, default=0,
                    help="receptor points per LDS tile (0: resident receptors where they fit)")
    ap.add_argument("--steps-per-dump", type=int, default=1000,
                    help="MC steps per launch window")
    ap.add_argument("--verbose", action="store_true")
    ap.add_argument("--autotune", type=int, default=0,
                    help="1: tune launch shapes first (times every candidate on the workload)")
    ap.add_argument("--tune-budget", type=float, default=90.0, help="seconds")
    ap.add_argument("--mode", default="da", choices=["plain", "da", "post"],
                    help="plain vs delayed acceptance (default), plain engine only, or post-MC")
    ap.add_argument("--no-record", action="store_true", help="plain mode: keep no records")
    ap.add_argument("--drain", action="store_true",
                    help="plain mode: copy every window's records to a host sink (counted), so "
                         "the rate includes the device-to-host transfers")
    ap.add_argument("--k-sub", type=int, default=4, help="surrogate steps per exact evaluation")
    ap.add_argument("--log-snapshots", type=int, default=32,
                    help="state snapshots taken over the warm-up for training")
    ap.add_argument("--train-poses", type=int, default=2000000)
    ap.add_argument("--epochs", type=int, default=20)
    ap.add_argument("--sur-mode", default="tab", choices=["tab", "direct"],
                    help="surrogate from tabulated network outputs or per-atom network")
    ap.add_argument("--batch-poses", type=int, default=4096)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--grid-spacing", type=float, default=0.5)
    ap.add_argument("--grid-half", type=float, default=14.0)
    ap.add_argument("--t-floor", type=float, default=None,
                    help="lowest kT of every ladder (default: the workload's 0.5)")
    ap.add_argument("--t-ceiling", type=float, default=None,
                    help="highest kT of every ladder (default: the workload's 5)")
    ap.add_argument("--trans-sigma", type=float, default=None,
                    help="translation step sigma, A (default: the workload's 0.2)")
    ap.add_argument("--rot-sigma-deg", type=float, default=None,
                    help="rotation step sigma, degrees (default: the workload's 3)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (HBM counter calibration)")
    args = ap.parse_args(argv)
    if args.mode == "da" and args.nc > 0 and args.nc % max(1, args.k_sub):
        ap.error(f"--nc {args.nc} must be a multiple of --k-sub {args.k_sub}: delayed "
                 "acceptance exchanges on whole groups of k sub-steps")
    return args


def round_up(n, m):
    return ((int(n) + m - 1) // m) * m if m > 0 else int(n)


def step_plan(args):
    """MC step counts of a run: (timed, warm-up, unit). Both are whole
    multiples of unit = lcm(k, nc) (k alone without exchanges), at least one
    unit when the bench step count is positive."""
    k = max(1, args.k_sub) if args.mode == "da" else 1
    unit = math.lcm(k, args.nc) if args.nc > 0 else k
    timed = round_up(max(args.steps, 0) * args.mc_per_step, unit)
    warm = round_up(max(args.warmup, 0) * args.mc_per_step, unit)
    return timed, warm, unit


def copy_probe(mib, reps=5):
    """Plain device-to-device copy of known size, for scaling FETCH_SIZE /
    WRITE_SIZE counters to bytes."""
    import torch
    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes_read": n * 4, "copy_probe_bytes_written": n * 4,
                      "reps": reps}))


def workload(args):
    """The complexes and their description; the regime flags (ladder, move
    sizes) override the workload's values and are named in the description."""
    from geaxdock_gpu_amd import synth
    kw, regime = {}, {}
    for flag, key in (("t_floor", "t_floor"), ("t_ceiling", "t_ceiling"),
                      ("trans_sigma", "trans_sigma"), ("rot_sigma_deg", "rot_sigma_deg")):
        v = getattr(args, flag, None)
        if v is not None:
            kw[key] = v
            regime[key] = v
    if args.config == "single":
        cxs = [synth.single_complex(seed=0, **kw)]
        desc = dict(workload="single complex: 1 lig x 1 prt x 8 temps, lna 32, pnp 1200")
    elif args.config == "batch":
        cxs = synth.uneven_batch(args.batch, seed=2, **kw)
        desc = dict(workload=f"uneven batch: {args.batch} complexes, lna U[12,64], pnp "
                             "U[800,4000], 5 lig x 3 prt x 8 temps each")
    else:
        cxs = [synth.ensemble(seed=1, **kw)]
        desc = dict(workload="ensemble: 20 lig x 11 prt x 16 temps = 3520 replicas, lna 48, "
                             "pnp 3000, pnk 200, mcs 8x40")
    if regime:
        desc["regime"] = regime
    return cxs, desc


def host_cpu():
    """CPU model and logical core count of this host (for the CPU baseline)."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"cpu_model": model, "logical_cpus": os.cpu_count()}


def cpu_baseline(cxs, args):
    """FP32 CPU path on a bounded sample of the workload: two ligand
    conformers x one protein conformer x all temperatures of the ensemble, or
    every replica of the first complex of a batch, for about cpu_seconds;
    the rate is scaled by pair count (lna * pnp) to the workload's mean."""
    import numpy as np
    from oracle import gdo
    from geaxdock_gpu_amd import synth
    if args.config == "ensemble":
        sub = synth.ensemble(seed=1, n_lig=2, n_prt=1)
    else:
        sub = cxs[0]
    d = sub.desc()
    st = gdo.McState(sub, rec_cap=0)
    threads = min(os.cpu_count() or 1, 16)
    gdo.mc_run(d, st, seed=0, init=True, step0=0, n_steps=0, nc=args.nc, threads=threads)
    chunk = max(args.nc, 20)
    steps = 0
    t0 = time.perf_counter()
    while True:
        gdo.mc_run(d, st, seed=0, init=False, step0=steps, n_steps=chunk, nc=args.nc,
                   threads=threads)
        steps += chunk
        dt = time.perf_counter() - t0
        if dt >= args.cpu_seconds:
            break
    rs = sub.n_rep * steps
    w_sub = sub.lna * sub.pnp
    w_job = np.mean([c.n_rep * c.lna * c.pnp for c in cxs]) / np.mean([c.n_rep for c in cxs])
    return {"value": rs / dt * (w_sub / w_job), "unit": "replica-steps/s", "cores": threads,
            "kind": "port", "host": host_cpu(),
            "sample": f"{sub.n_rep} replicas x {steps} MC steps (lna {sub.lna}, pnp {sub.pnp}); "
                      f"{dt:.1f} s on {threads} threads; scaled by pair count to the "
                      f"workload's mean of {w_job:.0f} pairs per replica-step",
            "one_core_single_complex": one_core_rate(args)}


def one_core_rate(args, seconds=3.0):
    """The same CPU path on one thread over the single-complex workload (1 lig x
    1 prt x 8 temperatures, lna 32, pnp 1200), for comparison with the
    reference's printed single-core figure of 843.75 MC sweeps per second
    (src/util.C:1357)."""
    from oracle import gdo
    from geaxdock_gpu_amd import synth
    cx = synth.single_complex(seed=0)
    d = cx.desc()
    st = gdo.McState(cx, rec_cap=0)
    gdo.mc_run(d, st, seed=0, init=True, step0=0, n_steps=0, nc=args.nc, threads=1)
    chunk = max(args.nc, 20)
    steps = 0
    t0 = time.perf_counter()
    while True:
        gdo.mc_run(d, st, seed=0, init=False, step0=steps, n_steps=chunk, nc=args.nc, threads=1)
        steps += chunk
        dt = time.perf_counter() - t0
        if dt >= seconds:
            break
    rate = cx.n_rep * steps / dt
    return {"value": rate, "unit": "replica-steps/s", "cores": 1,
            "sample": f"{cx.n_rep} replicas x {steps} steps, {dt:.1f} s",
            "reference_single_core": 843.75, "ratio_to_reference_single_core": rate / 843.75}


def kernel_names(eng, da=0):
    """Kernel instances the MC launch uses (one per atoms-per-lane class)."""
    apls = sorted({eng.shape(c)[1] for c in range(len(eng.complexes))})
    return ", ".join(f"gd::gd_mc_kernel<{a}, {da}>" for a in apls)


def progress(msg):
    """One line on stderr when a phase of the run ends."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def traffic_model(args):
    """The measured HBM traffic of this workload's plain launches, if profiled."""
    key = (args.config, args.batch if args.config == "batch" else None)
    if (args.config == "single" or args.tile or args.split > 0 or args.ladders > 0 or args.lpr
            or args.apl):
        return None
    return PLAIN_TRAFFIC.get(key)


def roofline(pairs_per_step, mc_steps, t_compute, launches, kname, traffic=None):
    """FP32 VALU roofline of an MC run: 36 flops per ligand-atom x
    receptor-point pair per replica-step, over the summed launch durations.
    traffic: PLAIN_TRAFFIC's model of the launches' HBM bytes, or None."""
    flops = FLOPS_PER_PAIR * pairs_per_step * mc_steps
    ach = flops / t_compute if t_compute > 0 else 0.0
    n = max(1, launches)
    out = {"bound": "valu", "unit": "TFLOP/s", "achieved": ach / 1e12,
           "peak": PEAK_FP32 / 1e12, "frac": ach / PEAK_FP32,
           "frac_unpacked_fp32": ach / PEAK_FP32_UNPACKED, "traffic": None,
           "kernel": kname, "launches": launches, "mc_steps": mc_steps,
           "kernel_ms_avg": t_compute * 1e3 / n,
           "flops_per_launch": flops / n, "flops": flops,
           "note": "FP32 vector (VALU) bound, not MFMA; 36 flops per ligand-receptor pair "
                   "per replica-step; peak is the packed-FP32 157.3 TFLOP/s, "
                   "frac_unpacked_fp32 uses 78.6"}
    if traffic is not None:
        rf, rs, wf, ws = traffic
        total = n * (rf + wf) + mc_steps * (rs + ws)
        out["traffic"] = total / n
        out["hbm_gbs"] = total / t_compute / 1e9 if t_compute > 0 else 0.0
        out["traffic_note"] = ("HBM bytes per launch from the workload's FETCH_SIZE/WRITE_SIZE "
                               "profile (%.3g B per launch + %.3g B per MC step), "
                               "copy-calibrated; far below 8 TB/s, so HBM is not the bound"
                               % (rf + wf, rs + ws))
    return out


def launch_options(args, cxs):
    """Launch shape: the autotuner's choice with --autotune 1, else the
    engine's automatic packing (whole ladders, then the receptor split that
    fills each workgroup to 8 waves), persistent for `batch`; explicit flags
    override."""
    from geaxdock_gpu_amd import abi
    if args.config == "batch":
        opts = dict(receptor_split=0, ladders_per_pack=0, launch=abi.LAUNCH_PERSISTENT)
    else:
        opts = dict(receptor_split=0, ladders_per_pack=0, launch=0)
    opts.update(lanes_per_replica=args.lpr, atoms_per_lane=args.apl, receptor_tile=args.tile)
    if args.split >= 0:
        opts["receptor_split"] = args.split
    if args.ladders >= 0:
        opts["ladders_per_pack"] = args.ladders
    if args.launch >= 0:
        opts["launch"] = args.launch
    tuned = None
    if args.autotune == 1:
        from geaxdock_gpu_amd import autotune
        t0 = time.perf_counter()
        best, trials = autotune.tune(cxs, nc=args.nc, budget_s=args.tune_budget)
        opts.update(best)
        tuned = {"best": best, "seconds": time.perf_counter() - t0,
                 "trials": [{**t["opts"], "ms_per_step": (t["s_per_step"] or 0) * 1e3}
                            for t in trials]}
    return opts, tuned


class BestLog:
    """Per-complex and per-ladder best energies of an engine's replicas."""

    def __init__(self, eng, cxs):
        self.eng, self.cxs = eng, cxs
        self.per_cx, self.per_ladder = {}, {}

    def take(self, tag):
        import numpy as np
        eng, cxs = self.eng, self.cxs
        bs = [eng.best(c)[0].astype(np.float64) for c in range(len(cxs))]
        lad = []
        for c, b in enumerate(bs):
            # a temperature ladder is one (protein conformer, ligand conformer)
            # pair; its best is the minimum over its replicas
            rep = eng.state(c)["replica"]
            key = rep[:, 1].astype(np.int64) * cxs[c].n_lig + rep[:, 3]
            keys, inv = np.unique(key, return_inverse=True)
            lb = np.full(len(keys), np.inf)
            np.minimum.at(lb, inv, b)
            lad.append(lb)
        self.per_cx[tag] = np.array([b.min() for b in bs])
        self.per_ladder[tag] = np.concatenate(lad)
        b = np.concatenate(bs)
        return {"best_energy": float(b.min()), "best_p01": float(np.percentile(b, 1)),
                "replica_best_mean": float(b.mean())}

    def delta(self, a, b):
        """Mean per-complex and per-ladder difference best[a] - best[b]."""
        import numpy as np
        dc = self.per_cx[a] - self.per_cx[b]
        dl = self.per_ladder[a] - self.per_ladder[b]
        return {"complex_mean": float(dc.mean()), "complex_max": float(dc.max()),
                "complex_min": float(dc.min()),
                "complexes_within_0p1": float(np.mean(np.abs(dc) <= 0.1)),
                "ladder_mean": float(dl.mean()),
                "ladders_within_0p1": float(np.mean(np.abs(dl) <= 0.1))}


def da_compare(args, cxs, desc):
    """Plain vs delayed acceptance on the same complexes, seed and start state."""
    import numpy as np
    import geaxdock_gpu_amd as gd
    from geaxdock_gpu_amd import surrogate as sg

    k = args.k_sub
    nc = args.nc
    mc_steps, mc_warm, unit = step_plan(args)
    n_rep = sum(c.n_rep for c in cxs)
    pairs_per_step = sum(c.n_rep * c.lna * c.pnp for c in cxs)
    opts, tuned = launch_options(args, cxs)
    dump = round_up(args.steps_per_dump, unit)

    # 1. warm-up, snapshotted for training
    eng = gd.Engine(seed=0, record=False, steps_per_dump=dump, **opts)
    for cx in cxs:
        eng.add_complex(cx)
    t0 = time.perf_counter()
    n_snap = max(1, min(args.log_snapshots, mc_warm // unit)) if mc_warm else 0
    every = round_up(max(1, mc_warm // max(1, n_snap)), unit) if n_snap else 0
    warm_kernel_s = 0.0
    recs, rcid = None, None
    if n_snap:
        recs, rcid, warm_kernel_s = sg.log_states(eng, n_snap, every, nc, with_time=True)
    rest = mc_warm - n_snap * every
    if rest > 0:
        warm_kernel_s += eng.run(rest, nc)["t_compute"]
    if recs is None:  # no warm-up: train on the start states
        recs, rcid = sg.log_states(eng, 1, unit, nc)
    t_log = time.perf_counter() - t0
    warm_steps = max(mc_warm, n_snap * every)
    progress(f"warm-up: {warm_steps} MC steps of {n_rep} replicas, {len(recs)} states logged "
             f"({t_log:.1f} s)")

    # 2. training set, training
    t0 = time.perf_counter()
    eng.sur_build_grids(args.grid_spacing, args.grid_half)
    ts = sg.build_training_set(eng, recs, rcid, args.train_poses, seed=1)
    m, sigma = sg.load_training_set(eng, ts)
    t_data = time.perf_counter() - t0
    progress(f"labelled {len(ts['t'])} training poses ({t_data:.1f} s)")
    fit = sg.fit(eng, ts, sigma, n_train=int(0.9 * len(ts["t"])), epochs=args.epochs,
                 batch_poses=args.batch_poses, lr=args.lr)
    progress(f"trained: held-out RMSE {fit['test_rmse_energy']:.4f} ({fit['seconds']:.1f} s)")
    # BF16 MFMA work of the training: per atom row and epoch the forward
    # (2 (64 x 128 + 128 x 128 + 128) flops) and backward (dA1 2 x 128 x 128,
    # dW2 2 x 128 x 128, dW1 2 x 64 x 128), plus the last epoch's evaluation
    # of the training range
    rows_train = float(ts["lna"][:fit["n_train"]].sum())
    train_flops = rows_train * (args.epochs * SUR_TRAIN_FLOPS_PER_ROW + SUR_FWD_FLOPS_PER_ROW)
    params = eng.sur_get_params()
    blob = eng.checkpoint()
    eng.close()
    del ts

    # 3. timed runs from one checkpoint
    run = gd.Engine(seed=0, record=False, steps_per_dump=dump, **opts)
    for cx in cxs:
        run.add_complex(cx)
    run.restore(blob)
    run.sur_build_grids(args.grid_spacing, args.grid_half)
    run.sur_set_params(params)
    t0 = time.perf_counter()
    if args.sur_mode == "tab":
        run.sur_tabulate()
    t_tab = time.perf_counter() - t0
    best = BestLog(run, cxs)

    def da_run(tag, scale):
        run.restore(blob)
        run.reset_best()
        t0 = time.perf_counter()
        st = run.run_da(mc_steps, nc, k, scale)
        dt = time.perf_counter() - t0
        evals = st["replica_steps"] / k
        r = {"mc_steps": mc_steps, "seconds": dt, "replica_steps_per_s": n_rep * mc_steps / dt,
             "kernel_replica_steps_per_s": n_rep * mc_steps / st["t_compute"],
             "exact_evals_per_s": evals / dt,
             "composite_acceptance": st["accepted"] / evals,
             "accepted_per_step": st["accepted"] / st["replica_steps"],
             "kernel_seconds": st["t_compute"], "launches": st["kernel_launches"]}
        r.update(best.take(tag))
        progress(f"{tag}: {mc_steps} MC steps, {r['replica_steps_per_s']:.4g} replica-steps/s, "
                 f"composite acceptance {r['composite_acceptance']:.3f}")
        return r

    out = {"da": da_run("da", sigma)}
    out["da_no_network"] = da_run("da_nonet", 0.0)

    # plain, in segments: best energies at the DA run's exact evaluations, at
    # its wall time (planned from the warm-up's kernel rate) and at equal steps
    t_da = out["da"]["seconds"]
    warm_rate = warm_steps / warm_kernel_s if warm_kernel_s > 0 else 0.0  # MC steps per second
    n_eval = round_up(mc_steps // k, nc if nc > 0 else 1)
    n_eq = round_up(t_da * warm_rate, nc if nc > 0 else 1) if warm_rate else n_eval
    at = {"plain_eq_eval": min(max(n_eval, nc), mc_steps),
          "plain_eq_time": min(max(n_eq, nc), mc_steps), "plain": mc_steps}
    run.restore(blob)
    run.reset_best()
    done, dt_plain, t_kernel, launches, accepted = 0, 0.0, 0.0, 0, 0
    seg = {}
    for mk in sorted(set(at.values())):
        t0 = time.perf_counter()
        st = run.run(mk - done, nc)
        dt_plain += time.perf_counter() - t0
        t_kernel += st["t_compute"]
        launches += st["kernel_launches"]
        accepted += st["accepted"]
        done = mk
        for tag, m in at.items():
            if m == mk:
                seg[tag] = dict(best.take(tag), mc_steps=mk, seconds=dt_plain)
    out["plain"] = dict(seg["plain"], replica_steps_per_s=n_rep * mc_steps / dt_plain,
                        kernel_replica_steps_per_s=n_rep * mc_steps / t_kernel,
                        acceptance=accepted / (n_rep * mc_steps), kernel_seconds=t_kernel,
                        launches=launches)
    out["plain_equal_exact_evals"] = seg["plain_eq_eval"]
    out["plain_equal_wall_time"] = seg["plain_eq_time"]
    progress(f"plain: {mc_steps} MC steps, {out['plain']['replica_steps_per_s']:.4g} "
             f"replica-steps/s")
    kname = kernel_names(run)
    kname_da = kernel_names(run, 2 if args.sur_mode == "tab" else 1)
    run.close()

    speed = out["da"]["replica_steps_per_s"] / out["plain"]["replica_steps_per_s"]
    speed_nonet = out["da_no_network"]["replica_steps_per_s"] / out["plain"]["replica_steps_per_s"]
    cmp = {"equal_steps": best.delta("da", "plain"),
           "equal_wall_time": best.delta("da", "plain_eq_time"),
           "equal_exact_evals": best.delta("da", "plain_eq_eval"),
           "no_network_equal_steps": best.delta("da_nonet", "plain"),
           "no_network_equal_exact_evals": best.delta("da_nonet", "plain_eq_eval"),
           "network_vs_no_network": best.delta("da", "da_nonet")}
    d_steps = cmp["equal_steps"]["complex_mean"]
    d_time = cmp["equal_wall_time"]["complex_mean"]
    roof = roofline(pairs_per_step, mc_steps, out["plain"]["kernel_seconds"],
                    out["plain"]["launches"], kname, traffic_model(args))
    # the DA kernel's exact share: one pose evaluation per k surrogate steps
    roof_da = roofline(pairs_per_step / k, mc_steps, out["da"]["kernel_seconds"],
                       out["da"]["launches"], kname_da)
    roof_da["note"] = "exact evaluations only (1 per k steps); surrogate sub-steps not counted"
    return {
        "metric": METRIC,
        "value": out["da"]["replica_steps_per_s"],
        "unit": "replica-steps/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": out["da"]["seconds"] / mc_steps * args.mc_per_step * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded generator, geaxdock_gpu_amd/synth.py)",
        "config": dict(desc, replicas=n_rep, mc_per_step=args.mc_per_step, mc_steps=mc_steps,
                       mc_warmup=warm_steps, steps_per_exchange=nc, k_sub=k,
                       surrogate=args.sur_mode, launch_window=dump),
        "roofline": roof,
        "roofline_da_exact_part": roof_da,
        "launch_options": opts, "autotune": tuned,
        "plain": out["plain"], "da": out["da"], "da_no_network": out["da_no_network"],
        "plain_equal_exact_evals": out["plain_equal_exact_evals"],
        "plain_equal_wall_time": out["plain_equal_wall_time"],
        "da_speedup": speed, "da_no_network_speedup": speed_nonet,
        "best_energy_delta": d_steps, "best_energy_delta_equal_wall_time": d_time,
        "best_energy_compare": cmp,
        "target_met": bool(speed >= 3.0 and abs(d_steps) <= 0.1 and d_time <= 0.1),
        "target": "DA >= 3x plain replica-steps/s; per-complex mean best energy within 0.1 of "
                  "plain at equal MC steps, and no worse than plain + 0.1 at equal wall time",
        "surrogate": {"log_seconds": t_log, "label_seconds": t_data, "sigma": sigma,
                      "tabulate_seconds": t_tab,
                      "train_poses": fit["n_train"], "test_poses": fit["n_test"],
                      "epochs": args.epochs, "train_seconds": fit["seconds"],
                      "train_loss": fit["train_loss"], "test_rmse_std": fit["test_rmse_std"],
                      "test_rmse_energy": fit["test_rmse_energy"],
                      "test_energy_std": fit["test_energy_std"],
                      "train_rows_per_epoch": rows_train,
                      "mfma": {"bound": "mfma", "unit": "TFLOP/s",
                               "achieved": train_flops / fit["seconds"] / 1e12,
                               "peak": PEAK_BF16 / 1e12,
                               "frac": train_flops / fit["seconds"] / PEAK_BF16,
                               "note": "BF16 MFMA work of the whole training call (forward "
                                       "49,408 + backward 81,920 flops per atom row and epoch, "
                                       "plus one evaluation pass) over its wall time, all "
                                       "kernels and launches included; peak: MI355X dense "
                                       "BF16 2.5 PFLOP/s"}},
    }


def post_bench(args, cxs, desc):
    """Post-MC clustering of the workload's accepted-step records: k-means
    medoids per replica (k 50), then over all of those (k 500), on the device.
    value = records clustered per second of wall time (grouping, the copy to
    HBM and both levels); the device-only rate and a roofline of the assign
    pass (27 flops per object-center distance) are reported beside it."""
    import numpy as np
    import geaxdock_gpu_amd as gd
    mc_steps, mc_warm, _ = step_plan(args)
    recs = []
    with gd.Engine(seed=0, record=True, steps_per_dump=2000) as eng:
        for cx in cxs:
            eng.add_complex(cx)
        if mc_warm > 0:
            eng.run(mc_warm, args.nc)
        eng.run(mc_steps, args.nc, sink=lambda c, r: recs.append(r))
        allr = np.concatenate(recs)
        eng.post_mc(allr, 50, 500, seed=1)  # warm-up
        tm = {}
        t0 = time.perf_counter()
        idx, sz = eng.post_mc(allr, 50, 500, seed=0, timing=tm)
        dt = time.perf_counter() - t0
        # level 1 alone for the roofline: records grouped by replica
        order = np.argsort(allr["replica"][:, 0], kind="stable")
        rep = allr["replica"][order, 0]
        counts = np.bincount(rep)
        counts = counts[counts > 0]
        off = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
        feat = allr["energy"]["e"][order, :9]
        t1 = {}
        _, _, passes = eng.kmeans_medoids(feat, off, 50, seed=0, timing=t1)
    live = counts >= 50
    obj_passes = float((counts[live] * passes[live]).sum())
    flops = obj_passes * 50 * 27.0
    achieved = flops / (t1["device_ms"] * 1e-3)
    out = {
        "metric": "post-MC records clustered/s (k-means medoids per replica, then over medoids)",
        "value": len(allr) / dt,
        "unit": "records/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (distances), int64 (cluster sums)",
        "data": "synthetic (seeded generator, geaxdock_gpu_amd/synth.py), records of a plain run",
        "config": dict(desc, records=int(len(allr)), replicas=int(len(counts)), k_replica=50,
                       k_final=500, medoids=int(len(idx)), mc_steps=mc_steps),
        "device_ms": tm["device_ms"],
        "device_records_per_s": len(allr) / (tm["device_ms"] * 1e-3),
        "roofline": {"bound": "valu", "achieved": achieved / 1e12, "peak": PEAK_FP32 / 1e12,
                     "unit": "TFLOP/s", "frac": achieved / PEAK_FP32, "traffic": None,
                     "kernel": "gd::gd_km_assign_kernel (+ update), level 1",
                     "work": f"{obj_passes:.3e} object-passes x 50 centers x 27 flops",
                     "passes_mean": float(passes[live].mean()), "passes_max": int(passes.max())},
    }
    if not args.no_cpu:
        try:
            out["cpu_baseline"] = post_cpu_baseline(allr, args)
        except Exception as ex:
            out["cpu_baseline"] = {"error": str(ex)}
    return out


def post_cpu_baseline(allr, args):
    """The oracle's post_mc (one thread) on the records of the first replicas,
    grown until about cpu_seconds of work."""
    import numpy as np
    from oracle import gdo
    reps = np.unique(allr["replica"][:, 0])
    n_r = 8
    while True:
        sub = allr[np.isin(allr["replica"][:, 0], reps[:n_r])]
        t0 = time.perf_counter()
        gdo.post_mc(sub, 50, 500, seed=0)
        dt = time.perf_counter() - t0
        if dt >= args.cpu_seconds or n_r >= len(reps):
            break
        n_r = min(len(reps), int(n_r * max(2.0, args.cpu_seconds / max(dt, 1e-3))))
    return {"value": len(sub) / dt, "unit": "records/s", "cores": 1, "kind": "port",
            "host": host_cpu(),
            "sample": f"{len(sub)} records of {n_r} replicas, {dt:.1f} s"}


def plain_bench(args, cxs, desc):
    """The plain engine alone: `steps` bench steps after the warm-up, records
    written to HBM as the reference's kernel does unless --no-record."""
    import geaxdock_gpu_amd as gd
    mc_steps, mc_warm, _ = step_plan(args)
    opts, tuned = launch_options(args, cxs)
    eng = gd.Engine(seed=0, record=not args.no_record, steps_per_dump=args.steps_per_dump,
                    verbose=args.verbose, **opts)
    t0 = time.perf_counter()
    for cx in cxs:
        eng.add_complex(cx)
    n_rep = sum(c.n_rep for c in cxs)
    if mc_warm > 0:
        eng.run(mc_warm, args.nc)
    t_setup = time.perf_counter() - t0  # host-to-device upload, packing, warm-up
    drained = [0, 0]

    def count(cid, recs):
        drained[0] += len(recs)
        drained[1] += recs.nbytes

    sink = count if (args.drain and not args.no_record) else None
    t0 = time.perf_counter()
    st = eng.run(mc_steps, args.nc, sink=sink)
    dt = time.perf_counter() - t0
    pairs_per_step = sum(c.n_rep * c.lna * c.pnp for c in cxs)
    roof = roofline(pairs_per_step, mc_steps, st["t_compute"], st["kernel_launches"],
                    kernel_names(eng), traffic_model(args))
    out = {
        "metric": "MC replica-steps/s (plain REMC engine)",
        "value": n_rep * mc_steps / dt,
        "unit": "replica-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt / mc_steps * args.mc_per_step * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded generator, geaxdock_gpu_amd/synth.py)",
        "config": dict(desc, replicas=n_rep, mc_per_step=args.mc_per_step, mc_steps=mc_steps,
                       mc_warmup=mc_warm, steps_per_exchange=args.nc,
                       launch_window=args.steps_per_dump, records=not args.no_record),
        "roofline": roof,
        "launch_options": opts,
        "autotune": tuned,
        "acceptance_ratio": st["accepted"] / max(1, st["replica_steps"]),
        "exchange_accepts": st["exchanges_accepted"],
        "t_compute_s": st["t_compute"],
        "kernel_replica_steps_per_s": n_rep * mc_steps / st["t_compute"],
        "setup_s": t_setup,
    }
    if sink is not None:
        out["drained"] = {"records": drained[0], "bytes": drained[1],
                          "note": "every window's records copied to the host inside the timed "
                                  "region (value includes the transfers)"}
    eng.close()
    return out


def main(argv=None):
    args = parse(argv)
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU per process; multi-GPU runs are "
                                   "not part of this build (shard whole temperature ladders or "
                                   "complexes across processes instead)"}))
        sys.exit(2)
    cxs, desc = workload(args)
    progress(f"workload: {desc['workload']}")
    if args.mode == "post":
        print(json.dumps(post_bench(args, cxs, desc)), flush=True)
        return
    out = da_compare(args, cxs, desc) if args.mode == "da" else plain_bench(args, cxs, desc)
    if not args.no_cpu:
        try:
            out["cpu_baseline"] = cpu_baseline(cxs, args)
        except Exception as ex:  # the GPU figures stand without it
            out["cpu_baseline"] = {"error": str(ex)}
        progress("CPU baseline done")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
