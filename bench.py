#!/usr/bin/env python3
"""Benchmark: MC replica-steps/s of the plain REMC engine on one MI355X.

Workload (default): the ensemble configuration -- one synthetic complex,
3,000-point receptor with 11 protein conformers, 48-atom ligand with 20
conformers, 16 temperatures (3,520 replicas), KDE 200, MCS 8 x 40, an exchange
attempt every 20 steps, FP32. A "step" is one MC step of every replica.

value = replicas * steps / seconds of the timed region (the steps run with all
inputs resident in HBM; accepted-state records are written to HBM as the
reference's kernel does, and not copied to the host).

roofline: algorithmic FP32 work of the fused MC kernel, 36 flops per
ligand-atom x receptor-point pair per replica-step (distance 10, soft LJ 13,
far-field electrostatics 5, PMF 8; conditional terms not counted), divided by
the kernel's mean duration from HIP events on the engine's stream; peak is
MI355X FP32 157.3 TFLOP/s.

cpu_baseline: the FP32 CPU restatement (oracle/, OpenMP over replicas) on a
bounded sample of the same workload, on this host's cores.
"""
import argparse
import json
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
    ap.add_argument("--cpu-sample-steps", type=int, default=30)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--lpr", type=int, default=0)
    ap.add_argument("--apl", type=int, default=0)
    ap.add_argument("--split", type=int, default=0, help="receptor split (0 = auto, 1, 2)")
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
        cxs = [synth.config0(seed=0)]
        desc = dict(workload="single complex: 1 lig x 1 prt x 8 temps, lna 32, pnp 1200")
    elif args.config == "batch":
        cxs = synth.config_batch(args.batch, seed=2)
        desc = dict(workload=f"uneven batch: {args.batch} complexes, lna U[12,64], pnp U[800,4000], "
                             "5 lig x 3 prt x 8 temps each")
    else:
        cxs = [synth.config1(seed=1)]
        desc = dict(workload="ensemble: 20 lig x 11 prt x 16 temps = 3520 replicas, lna 48, "
                             "pnp 3000, pnk 200, mcs 8x40")
    return cxs, desc


def cpu_baseline(cxs, args):
    """FP32 CPU path on a bounded sample: two ligand conformers x one protein
    conformer x all temperatures of the first complex."""
    import numpy as np
    from oracle import gdo
    from geaxdock_gpu_amd import synth
    cx0 = cxs[0]
    if args.config == "ensemble":
        sub = synth.config1(seed=1, n_lig=2, n_prt=1)
    elif args.config == "single":
        sub = cx0
    else:
        sub = cx0
    d = sub.desc()
    st = gdo.McState(sub, rec_cap=0)
    threads = os.cpu_count() or 1
    threads = min(threads, 16)
    gdo.mc_run(d, st, seed=0, init=True, step0=0, n_steps=0, nc=args.nc, threads=threads)
    steps = args.cpu_sample_steps
    t0 = time.perf_counter()
    gdo.mc_run(d, st, seed=0, init=False, step0=0, n_steps=steps, nc=args.nc, threads=threads)
    dt = time.perf_counter() - t0
    rs = sub.n_rep * steps
    # per replica-step cost scales with lna * pnp; rescale to the workload's mean
    w_sub = sub.lna * sub.pnp
    w_job = np.mean([c.lna * c.pnp for c in cxs])
    return {"value": rs / dt * (w_sub / w_job), "unit": "replica-steps/s", "cores": threads,
            "kind": "port",
            "sample": f"{sub.n_rep} replicas x {steps} steps (lna {sub.lna}, pnp {sub.pnp}); "
                      f"{dt:.1f} s on {threads} threads; scaled by pair count to the workload"}


def kernel_names(eng):
    """Kernel instances the MC launch uses (one per atoms-per-lane class)."""
    apls = sorted({eng.shape(c)[1] for c in range(len(eng.complexes))})
    return ", ".join(f"gd::gd_mc_kernel<{a}, false>" for a in apls)


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
    eng = gd.Engine(seed=0, record=True, steps_per_dump=2000, lanes_per_replica=args.lpr,
                    atoms_per_lane=args.apl, receptor_split=args.split)
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
    steps_per_launch = min(args.steps, 2000)
    kernel_s = st["kernel_ms_avg"] * 1e-3
    flops_launch = FLOPS_PER_PAIR * pairs_per_step * steps_per_launch
    achieved = flops_launch / kernel_s if kernel_s > 0 else 0.0
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
        "roofline": {"bound": "mfma", "unit": "TFLOP/s", "achieved": achieved / 1e12,
                     "peak": PEAK_FP32 / 1e12, "frac": achieved / PEAK_FP32, "traffic": None,
                     "kernel": kernel_names(eng), "kernel_ms_avg": st["kernel_ms_avg"],
                     "flops_per_launch": flops_launch,
                     "note": "FP32 vector (VALU) bound; 36 flops per ligand-receptor pair"},
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
