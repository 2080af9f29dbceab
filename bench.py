#!/usr/bin/env python3
"""Benchmark: MC replica-steps/s, plain vs surrogate delayed acceptance, on one MI355X.

Workload (default): the ensemble configuration -- one synthetic complex,
3,000-point receptor with 11 protein conformers, 48-atom ligand with 20
conformers, 16 temperatures (3,520 replicas), KDE 200, MCS 8 x 40, an exchange
attempt every 20 steps, FP32. A "step" is one MC step of every replica.

Default mode (da): a plain warm-up, 1,000 logged plain steps, surrogate
training on the GPU, then from one checkpoint `steps` plain steps and `steps`
delayed-acceptance steps (k surrogate steps per exact evaluation), both with
all inputs resident in HBM and no records. value = the DA run's replica-steps
per second of wall time; the plain figure, the speed-up and the best total
energy each run reached are reported beside it, with a DA run of equal wall
time and a DA run with the network removed (scale 0). --mode plain times the plain
engine alone (records written to HBM as the reference's kernel does).

roofline: algorithmic FP32 work of the plain MC kernel, 36 flops per
ligand-atom x receptor-point pair per replica-step (distance 10, soft LJ 13,
far-field electrostatics 5, PMF 8; conditional terms not counted), divided by
the kernel's mean duration from HIP events on the engine's stream; peak is
MI355X FP32 157.3 TFLOP/s.

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


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--nc", type=int, default=20, help="steps per exchange")
    ap.add_argument("--config", default="ensemble", choices=["ensemble", "single", "batch"])
    ap.add_argument("--batch", type=int, default=256, help="complexes in the batch workload")
    ap.add_argument("--cpu-seconds", type=float, default=12.0,
                    help="length of the CPU-path sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--lpr", type=int, default=0)
    ap.add_argument("--apl", type=int, default=0)
    ap.add_argument("--split", type=int, default=0, help="receptor split (0 = auto, 1, 2)")
    ap.add_argument("--ladders", type=int, default=0, help="temperature ladders per workgroup")
    ap.add_argument("--verbose", action="store_true")
    ap.add_argument("--launch", type=int, default=0, help="GD_LAUNCH_* flags")
    ap.add_argument("--autotune", type=int, default=-1,
                    help="tune launch shapes first (default: on for the batch workload)")
    ap.add_argument("--tune-budget", type=float, default=90.0, help="seconds")
    ap.add_argument("--mode", default="da", choices=["plain", "da", "post"],
                    help="plain vs surrogate delayed acceptance (default), or plain engine only")
    ap.add_argument("--no-record", action="store_true", help="plain mode: keep no records")
    ap.add_argument("--k-sub", type=int, default=4, help="surrogate steps per exact evaluation")
    ap.add_argument("--log-steps", type=int, default=1000, help="plain steps logged for training")
    ap.add_argument("--log-snapshots", type=int, default=50,
                    help="state snapshots taken over the logged steps")
    ap.add_argument("--train-poses", type=int, default=200000)
    ap.add_argument("--epochs", type=int, default=20)
    ap.add_argument("--sur-mode", default="tab", choices=["tab", "direct"],
                    help="surrogate from tabulated network outputs or per-atom network")
    ap.add_argument("--batch-poses", type=int, default=4096)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--grid-spacing", type=float, default=0.5)
    ap.add_argument("--grid-half", type=float, default=14.0)
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (HBM counter calibration)")
    return ap.parse_args()


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
    from geaxdock_gpu_amd import synth
    if args.config == "single":
        cxs = [synth.single_complex(seed=0)]
        desc = dict(workload="single complex: 1 lig x 1 prt x 8 temps, lna 32, pnp 1200")
    elif args.config == "batch":
        cxs = synth.uneven_batch(args.batch, seed=2)
        desc = dict(workload=f"uneven batch: {args.batch} complexes, lna U[12,64], pnp U[800,4000], "
                             "5 lig x 3 prt x 8 temps each")
    else:
        cxs = [synth.ensemble(seed=1)]
        desc = dict(workload="ensemble: 20 lig x 11 prt x 16 temps = 3520 replicas, lna 48, "
                             "pnp 3000, pnk 200, mcs 8x40")
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
    """FP32 CPU path on a bounded sample: two ligand conformers x one protein
    conformer x all temperatures of the first complex."""
    import numpy as np
    from oracle import gdo
    from geaxdock_gpu_amd import synth
    cx0 = cxs[0]
    if args.config == "ensemble":
        sub = synth.ensemble(seed=1, n_lig=2, n_prt=1)
    elif args.config == "single":
        sub = cx0
    else:
        sub = cx0
    d = sub.desc()
    st = gdo.McState(sub, rec_cap=0)
    threads = os.cpu_count() or 1
    threads = min(threads, 16)
    gdo.mc_run(d, st, seed=0, init=True, step0=0, n_steps=0, nc=args.nc, threads=threads)
    # windows of exchange segments until about cpu_seconds of work
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
    # per replica-step cost scales with lna * pnp; rescale to the workload's mean
    w_sub = sub.lna * sub.pnp
    w_job = np.mean([c.lna * c.pnp for c in cxs])
    return {"value": rs / dt * (w_sub / w_job), "unit": "replica-steps/s", "cores": threads,
            "kind": "port", "host": host_cpu(),
            "sample": f"{sub.n_rep} replicas x {steps} steps (lna {sub.lna}, pnp {sub.pnp}); "
                      f"{dt:.1f} s on {threads} threads; scaled by pair count to the workload",
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


def kernel_names(eng):
    """Kernel instances the MC launch uses (one per atoms-per-lane class)."""
    apls = sorted({eng.shape(c)[1] for c in range(len(eng.complexes))})
    return ", ".join(f"gd::gd_mc_kernel<{a}, 0>" for a in apls)


def progress(msg):
    """One line on stderr when a phase of the run ends."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def da_compare(args, cxs, desc):
    """Plain vs delayed acceptance on the same complexes, seed and start state.

    1. plain warmup, then snapshots of the replicas' states during a plain run;
    2. label training poses (logged states and states moved by 1-4
       proposals) with the exact energies, train the MLP on the GPU, report
       the held-out RMSE;
    3. from one checkpoint, time `steps` plain steps and `steps` DA steps
       (k_sub surrogate steps per exact evaluation) without records, and
       compare replica-steps/s and the best total energy reached.
    """
    import numpy as np
    import geaxdock_gpu_amd as gd
    from geaxdock_gpu_amd import surrogate as sg

    k = args.k_sub
    if args.steps % k or args.warmup % k or args.log_steps % k or args.nc % k:
        raise SystemExit("steps, warmup, log-steps and nc must be multiples of k-sub")
    n_rep = sum(c.n_rep for c in cxs)
    opts, tuned = launch_options(args, cxs)
    eng = gd.Engine(seed=0, record=False, steps_per_dump=2000, **opts)
    for cx in cxs:
        eng.add_complex(cx)
    eng.run(args.warmup, args.nc)
    progress(f"plain warm-up: {args.warmup} steps of {n_rep} replicas")
    t0 = time.perf_counter()
    # state snapshots every log_every steps over log_steps plain steps
    every = max(args.nc, args.log_steps // max(1, args.log_snapshots))
    recs, rcid = sg.log_states(eng, max(1, args.log_steps // every), every, args.nc)
    t_log = time.perf_counter() - t0
    progress(f"logged {len(recs)} replica states")
    t0 = time.perf_counter()
    eng.sur_build_grids(args.grid_spacing, args.grid_half)
    ts = sg.build_training_set(eng, recs, rcid, args.train_poses, seed=1)
    m, sigma = sg.load_training_set(eng, ts)
    t_data = time.perf_counter() - t0
    progress(f"labelled {len(ts['t'])} training poses")
    fit = sg.fit(eng, ts, sigma, n_train=int(0.9 * len(ts["t"])), epochs=args.epochs,
                 batch_poses=args.batch_poses, lr=args.lr)
    progress(f"trained: held-out RMSE {fit['test_rmse_energy']:.4f}")
    params = eng.sur_get_params()
    blob = eng.checkpoint()
    eng.close()

    run = gd.Engine(seed=0, record=False, steps_per_dump=2000, **opts)
    for cx in cxs:
        run.add_complex(cx)
    run.restore(blob)
    run.sur_build_grids(args.grid_spacing, args.grid_half)
    run.sur_set_params(params)
    t0 = time.perf_counter()
    if args.sur_mode == "tab":
        run.sur_tabulate()
    t_tab = time.perf_counter() - t0

    per_cx, per_ladder = {}, {}

    def ladder_best(c, b):
        # a temperature ladder is one (protein conformer, ligand conformer)
        # pair; its best is the minimum over its replicas
        rep = run.state(c)["replica"]
        key = rep[:, 1].astype(np.int64) * cxs[c].n_lig + rep[:, 3]
        keys, inv = np.unique(key, return_inverse=True)
        lb = np.full(len(keys), np.inf)
        np.minimum.at(lb, inv, b)
        return lb

    def best_stats(tag):
        bs = [run.best(c)[0].astype(np.float64) for c in range(len(cxs))]
        per_cx[tag] = np.array([b.min() for b in bs])
        per_ladder[tag] = np.concatenate([ladder_best(c, b) for c, b in enumerate(bs)])
        b = np.concatenate(bs)
        return {"best_energy": float(b.min()), "best_p01": float(np.percentile(b, 1)),
                "replica_best_mean": float(b.mean())}

    def one(mode, steps, tag, scale=None):
        run.restore(blob)
        run.reset_best()
        t0 = time.perf_counter()
        if mode == "plain":
            st = run.run(steps, args.nc)
        else:
            st = run.run_da(steps, args.nc, k, sigma if scale is None else scale)
        dt = time.perf_counter() - t0
        r = {"steps": steps, "replica_steps_per_s": n_rep * steps / dt,
             "kernel_replica_steps_per_s": n_rep * steps / st["t_compute"],
             "accepted_per_exact_eval": st["accepted"] / (st["replica_steps"] / (k if mode == "da" else 1)),
             "accepted_per_step": st["accepted"] / st["replica_steps"],
             "kernel_ms_avg": st["kernel_ms_avg"], "seconds": dt}
        r.update(best_stats(tag))
        progress(f"{tag}: {steps} steps, {r['replica_steps_per_s']:.4g} replica-steps/s")
        return r

    out = {"plain": one("plain", args.steps, "plain"), "da": one("da", args.steps, "da")}
    # the same wall-clock budget as the plain run
    eq = int(args.steps * out["da"]["replica_steps_per_s"] / out["plain"]["replica_steps_per_s"])
    step = math.lcm(k, args.nc) if args.nc > 0 else k
    eq = max(step, eq // step * step)
    out["da_equal_time"] = one("da", eq, "da_eq")
    # the same delayed acceptance with the network removed (scale 0: the
    # screen keeps only the exact cheap terms), for what the network adds
    out["da_no_network"] = one("da", args.steps, "da_nonet", scale=0.0)
    kname = kernel_names(run)
    run.close()
    speed = out["da"]["replica_steps_per_s"] / out["plain"]["replica_steps_per_s"]
    # best-pose energy per complex (energies of different complexes do not compare)
    dcx = per_cx["da"] - per_cx["plain"]
    d_best = float(dcx.mean()) if len(cxs) > 1 else float(dcx[0])
    best_cmp = {"complexes": len(cxs), "delta_mean": float(dcx.mean()),
                "delta_max": float(dcx.max()), "delta_min": float(dcx.min()),
                "within_0p1": float(np.mean(np.abs(dcx) <= 0.1)),
                "delta_equal_time_mean": float((per_cx["da_eq"] - per_cx["plain"]).mean()),
                "delta_no_network_mean": float((per_cx["da_nonet"] - per_cx["plain"]).mean())}
    # the same comparison per temperature ladder: many more samples than the
    # one minimum per complex, so a steadier view of what DA gives up
    dl = per_ladder["da"] - per_ladder["plain"]
    best_cmp.update({"ladders": int(len(dl)), "ladder_delta_mean": float(dl.mean()),
                     "ladders_within_0p1": float(np.mean(np.abs(dl) <= 0.1)),
                     "ladder_delta_equal_time_mean":
                         float((per_ladder["da_eq"] - per_ladder["plain"]).mean()),
                     "ladder_delta_no_network_mean":
                         float((per_ladder["da_nonet"] - per_ladder["plain"]).mean())})
    pairs_per_step = sum(c.n_rep * c.lna * c.pnp for c in cxs)
    roof = roofline(pairs_per_step, min(args.steps, 2000), out["plain"]["kernel_ms_avg"], kname,
                    desc)
    # the DA kernel's exact work: one pose evaluation per k surrogate steps
    roof_da = roofline(pairs_per_step / k, min(args.steps, 2000), out["da"]["kernel_ms_avg"],
                       kname.replace(", 0>", ", 2>"), None)
    return {
        "metric": "MC replica-steps/s, plain vs surrogate-DA, at matched best-pose energy",
        "value": out["da"]["replica_steps_per_s"],
        "unit": "replica-steps/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": out["da"]["seconds"] / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded generator, geaxdock_gpu_amd/synth.py)",
        "config": dict(desc, replicas=n_rep, steps_per_exchange=args.nc, k_sub=k,
                       surrogate=args.sur_mode),
        "roofline": roof,
        "roofline_da_exact_part": roof_da,
        "launch_options": opts, "autotune": tuned,
        "plain": out["plain"], "da": out["da"], "da_equal_time": out["da_equal_time"],
        "da_no_network": out["da_no_network"],
        "da_speedup": speed, "best_energy_delta": d_best, "best_energy_per_complex": best_cmp,
        "target_met": bool(speed >= 3.0 and abs(d_best) <= 0.1),
        "surrogate": {"log_seconds": t_log, "label_seconds": t_data, "sigma": sigma,
                      "tabulate_seconds": t_tab,
                      "train_poses": fit["n_train"], "test_poses": fit["n_test"],
                      "epochs": args.epochs, "train_seconds": fit["seconds"],
                      "train_loss": fit["train_loss"], "test_rmse_std": fit["test_rmse_std"],
                      "test_rmse_energy": fit["test_rmse_energy"],
                      "test_energy_std": fit["test_energy_std"]},
    }


# HBM bytes of one 2,000-step plain MC launch on the ensemble workload without
# records: rocprofv3 FETCH_SIZE 3845.0 KiB x 2 + WRITE_SIZE 786.8 KiB x 1 (the
# factors a plain 256 MiB device copy gives), from
# profiles/gd_mc_kernel_apl3_ensemble_3520rep_traffic.md
TRAFFIC_ENSEMBLE_2000 = (3845.0 * 2 + 786.8125) * 1024


def roofline(pairs_per_step, steps_per_launch, kernel_ms, kname, desc):
    """FP32 VALU roofline of an MC launch: 36 flops per ligand-atom x
    receptor-point pair, divided by the launch's HIP-event duration."""
    flops = FLOPS_PER_PAIR * pairs_per_step * steps_per_launch
    ach = flops / (kernel_ms * 1e-3) if kernel_ms > 0 else 0.0
    traffic = None
    if desc is not None and desc.get("workload", "").startswith("ensemble") and \
            steps_per_launch == 2000:
        traffic = TRAFFIC_ENSEMBLE_2000
    return {"bound": "mfma", "unit": "TFLOP/s", "achieved": ach / 1e12, "peak": PEAK_FP32 / 1e12,
            "frac": ach / PEAK_FP32, "frac_unpacked_fp32": ach / 78.6e12, "traffic": traffic,
            "kernel": kname, "kernel_ms_avg": kernel_ms, "flops_per_launch": flops,
            "note": "FP32 vector (VALU) bound; 36 flops per ligand-receptor pair; peak is the "
                    "packed-FP32 157.3 TFLOP/s, frac_unpacked_fp32 uses 78.6"}


def launch_options(args, cxs):
    opts = dict(lanes_per_replica=args.lpr, atoms_per_lane=args.apl, receptor_split=args.split,
                ladders_per_pack=args.ladders, launch=args.launch)
    tuned = None
    if args.autotune == 1 or (args.autotune == -1 and args.config == "batch"):
        from geaxdock_gpu_amd import autotune
        t0 = time.perf_counter()
        best, trials = autotune.tune(cxs, nc=args.nc, budget_s=args.tune_budget)
        opts.update(best)
        tuned = {"best": best, "seconds": time.perf_counter() - t0,
                 "trials": [{**t["opts"], "ms_per_step": (t["s_per_step"] or 0) * 1e3}
                            for t in trials]}
    return opts, tuned


def post_bench(args, cxs, desc):
    """Post-MC clustering of the workload's accepted-step records: k-means
    medoids per replica (k 50), then over all of those (k 500), on the device.
    value = records clustered per second of wall time (grouping, the copy to
    HBM and both levels); the device-only rate and a roofline of the assign
    pass (27 flops per object-center distance) are reported beside it."""
    import numpy as np
    import geaxdock_gpu_amd as gd
    from geaxdock_gpu_amd import abi
    recs = []
    with gd.Engine(seed=0, record=True, steps_per_dump=2000) as eng:
        for cx in cxs:
            eng.add_complex(cx)
        if args.warmup > 0:
            eng.run(args.warmup, args.nc)
        eng.run(args.steps, args.nc, sink=lambda c, r: recs.append(r))
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
                       k_final=500, medoids=int(len(idx))),
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


def main():
    args = parse()
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU per process; multi-GPU runs are "
                                   "not part of this build (shard whole temperature ladders or "
                                   "complexes across processes instead)"}))
        sys.exit(2)
    import numpy as np
    import geaxdock_gpu_amd as gd

    cxs, desc = workload(args)
    if args.mode == "post":
        print(json.dumps(post_bench(args, cxs, desc)))
        return
    if args.mode == "da":
        out = da_compare(args, cxs, desc)
        if not args.no_cpu:
            try:
                out["cpu_baseline"] = cpu_baseline(cxs, args)
            except Exception as ex:  # the GPU figures stand without it
                out["cpu_baseline"] = {"error": str(ex)}
        print(json.dumps(out))
        return
    opts, tuned = launch_options(args, cxs)
    eng = gd.Engine(seed=0, record=not args.no_record, steps_per_dump=2000, verbose=args.verbose,
                    **opts)
    for cx in cxs:
        eng.add_complex(cx)
    n_rep = sum(c.n_rep for c in cxs)
    if args.warmup > 0:
        eng.run(args.warmup, args.nc)
    t0 = time.perf_counter()
    st = eng.run(args.steps, args.nc)
    dt = time.perf_counter() - t0
    value = n_rep * args.steps / dt
    pairs_per_step = sum(c.n_rep * c.lna * c.pnp for c in cxs)
    # dominant kernel: one launch = min(steps, 2000) steps of every replica
    roof = roofline(pairs_per_step, min(args.steps, 2000), st["kernel_ms_avg"], kernel_names(eng),
                    desc if args.no_record else None)
    state_accept = st["accepted"] / max(1, st["replica_steps"])
    out = {
        "metric": "MC replica-steps/s (plain REMC engine)",
        "value": value,
        "unit": "replica-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded generator, geaxdock_gpu_amd/synth.py)",
        "config": dict(desc, replicas=n_rep, steps_per_exchange=args.nc),
        "roofline": roof,
        "launch_options": opts,
        "autotune": tuned,
        "acceptance_ratio": state_accept,
        "exchange_accepts": st["exchanges_accepted"],
        "t_compute_s": st["t_compute"],
    }
    eng.close()
    if not args.no_cpu:
        try:
            out["cpu_baseline"] = cpu_baseline(cxs, args)
        except Exception as ex:  # the GPU figure stands without it
            out["cpu_baseline"] = {"error": str(ex)}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
