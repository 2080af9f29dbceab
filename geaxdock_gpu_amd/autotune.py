"""Launch-shape autotuner for a set of complexes.

The engine's launch has four knobs that trade occupancy, lane utilization
and load balance against each other, and the best setting depends on the
mix of ligand and receptor sizes:

  receptor_split    0 (per workgroup: whole ladders, then the split that
                    fills 8 waves), or 1 or 2 lane groups per ladder for every
                    workgroup, each over its share of every receptor chunk;
  ladders_per_pack  temperature ladders per workgroup (replicas per
                    workgroup = ladders x n_tmp);
  launch            persistent workgroups or one workgroup per pack;
  lanes_per_replica lanes per replica (8..64; 0 lets the engine choose per
                    complex from its atom count);
  receptor_tile     receptor points per LDS tile (0: each receptor resident in
                    LDS when it fits, tiles of 256 otherwise; 64..512: every
                    receptor streams through LDS tiles of that length).

`tune` times each candidate on the real complexes for a few MC steps (a
fresh engine per candidate, same seed, no records) and returns the fastest
options: first the shape knobs with the default tiling, then the tile
lengths with the best shape. Results are cached in a JSON file keyed by the
workload's shape.
"""
import hashlib
import json
import os
import time

from . import abi


def signature(cxs, nc):
    h = hashlib.sha1()
    for cx in cxs:
        h.update(f"{cx.lna},{cx.pnp},{cx.pnk},{cx.n_lig},{cx.n_prt},{cx.n_tmp};".encode())
    h.update(f"nc={nc}".encode())
    return h.hexdigest()[:16]


TILES = (64, 128, 256, 512)


def candidates(cxs, full=False):
    """Shape option sets to try (default tiling). The default grid covers the
    knobs that matter on uneven batches; `full` adds the lane-shape axis."""
    out = []
    for split in (0, 1, 2):
        for ladders in (0, 1, 2, 3, 4):
            for launch in (0, abi.LAUNCH_PERSISTENT):
                out.append(dict(receptor_split=split, ladders_per_pack=ladders, launch=launch,
                                lanes_per_replica=0, receptor_tile=0))
    if full:
        for lpr in (8, 16, 32):
            out.append(dict(receptor_split=0, ladders_per_pack=0, launch=0,
                            lanes_per_replica=lpr, receptor_tile=0))
    return out


def tile_candidates(best, tiles=TILES):
    """The best shape with each receptor tile length."""
    return [dict(best, receptor_tile=t) for t in tiles]


def time_candidate(cxs, opts, steps=40, warmup=8, nc=20, seed=0):
    """Kernel seconds per MC step of the batch under `opts` (None if the
    engine rejects the shape)."""
    import geaxdock_gpu_amd as gd
    try:
        with gd.Engine(seed=seed, record=False, steps_per_dump=max(steps, warmup), **opts) as eng:
            for cx in cxs:
                eng.add_complex(cx)
            eng.run(warmup, nc)
            st = eng.run(steps, nc)
            return st["t_compute"] / steps
    except gd.GeauxDockError:
        return None


def tune(cxs, nc=20, steps=40, warmup=8, cache=None, full=False, budget_s=120.0, log=None,
         tiles=TILES, shapes=None):
    """Fastest option set for these complexes: (options, table of trials).
    `shapes` replaces the shape grid of the first stage; `tiles` are the
    tile lengths of the second."""
    key = signature(cxs, nc)
    if cache and os.path.exists(cache):
        with open(cache) as f:
            db = json.load(f)
        if key in db:
            return db[key]["best"], db[key]["trials"]
    t_end = time.perf_counter() + budget_s
    trials = []

    def run(cands):
        for opts in cands:
            if time.perf_counter() > t_end:
                break
            s = time_candidate(cxs, opts, steps, warmup, nc)
            trials.append({"opts": opts, "s_per_step": s})
            if log:
                log(f"autotune {opts} -> {s}")

    def fastest():
        ok = [t for t in trials if t["s_per_step"] is not None]
        if not ok:
            raise RuntimeError("autotune: no candidate launch shape was accepted")
        return min(ok, key=lambda t: t["s_per_step"])["opts"]

    run(shapes if shapes is not None else candidates(cxs, full))
    run(tile_candidates(fastest(), tiles))
    best = fastest()
    if cache:
        db = {}
        if os.path.exists(cache):
            with open(cache) as f:
                db = json.load(f)
        db[key] = {"best": best, "trials": trials}
        with open(cache, "w") as f:
            json.dump(db, f, indent=1)
    return best, trials
