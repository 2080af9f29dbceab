"""bench.py's command line as the benchmark driver calls it (no GPU needed)."""
import json
import math
import os
import subprocess
import sys

import bench

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_driver_flags_give_whole_da_groups():
    """`--gpus 1 --steps 20 --warmup 5`: 5,000 timed MC steps (the batch
    workload's run length) and a warm-up rounded up to whole groups of
    lcm(k, nc) steps, so the delayed-acceptance run accepts every count."""
    a = bench.parse(["--gpus", "1", "--steps", "20", "--warmup", "5"])
    assert a.config == "batch" and a.batch == 1024 and a.mode == "da"
    timed, warm, unit = bench.step_plan(a)
    assert unit == math.lcm(a.k_sub, a.nc) == 20
    assert timed == 5000 and warm == 1260
    assert timed % a.k_sub == 0 and warm % a.k_sub == 0 and a.nc % a.k_sub == 0
    assert bench.round_up(a.steps_per_dump, unit) % a.k_sub == 0


def test_any_step_counts_round_to_valid_groups():
    for steps in (1, 3, 7, 20, 33):
        for warm in (0, 1, 5, 9):
            for k in (1, 2, 4, 5, 10):
                a = bench.parse(["--steps", str(steps), "--warmup", str(warm), "--k-sub", str(k)])
                timed, w, unit = bench.step_plan(a)
                assert timed >= steps * a.mc_per_step and timed % unit == 0
                assert w >= warm * a.mc_per_step and w % unit == 0
                assert unit % k == 0 and unit % a.nc == 0


def test_plain_mode_needs_no_k_rounding():
    a = bench.parse(["--mode", "plain", "--steps", "3", "--warmup", "1", "--nc", "0"])
    timed, warm, unit = bench.step_plan(a)
    assert (timed, warm, unit) == (750, 250, 1)


def test_more_than_one_gpu_is_refused_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in json.loads(r.stdout.strip().splitlines()[-1])["error"]


def test_exchange_interval_must_hold_whole_groups():
    import pytest
    with pytest.raises(SystemExit):
        bench.parse(["--k-sub", "8"])
    a = bench.parse(["--k-sub", "8", "--nc", "40"])
    timed, warm, unit = bench.step_plan(a)
    assert unit == 40 and timed % 40 == 0


def test_roofline_traffic_only_for_profiled_workloads():
    """The roofline's HBM traffic comes from the counter profile of the same
    workload and launch shape; any other workload reports null."""
    a = bench.parse(["--gpus", "1", "--steps", "20", "--warmup", "5"])
    model = bench.traffic_model(a)
    assert model is not None
    rf, rs, wf, ws = model
    # 7 windows of 5,000 steps: per launch = the fixed part + steps/launch x per step
    r = bench.roofline(91744 * 122880, 5000, 60.0, 7, "k", model)
    assert math.isclose(r["traffic"], rf + wf + 5000 / 7 * (rs + ws), rel_tol=1e-12)
    assert math.isclose(r["hbm_gbs"], 7 * r["traffic"] / 60.0 / 1e9, rel_tol=1e-12)
    assert r["bound"] == "valu" and r["hbm_gbs"] < 1.0  # GB/s, against 8 TB/s
    for argv in (["--config", "batch", "--batch", "256"], ["--config", "single"],
                 ["--split", "2"], ["--tile", "256"], ["--ladders", "1"], ["--lpr", "16"],
                 ["--apl", "2"]):
        assert bench.traffic_model(bench.parse(argv)) is None
        assert bench.roofline(1, 1, 1.0, 1, "k", None)["traffic"] is None
    assert bench.traffic_model(bench.parse(["--config", "ensemble"])) is not None
