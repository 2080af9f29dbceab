"""The drop-in C++ Run() (include/geauxdock_run.hpp) on the reference's own
structure layout gives the same records as the engine's C API, and writes the
reference's printStates trajectory format."""
import os
import subprocess

import numpy as np
import pytest

from geaxdock_gpu_amd import abi, synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "run_ref_test")


def write_input(cx, path, steps, dump, nc, exchange, seed):
    i32 = lambda *v: np.array(v, np.int32).tobytes()
    with open(path, "wb") as f:
        f.write(i32(cx.n_lig, cx.n_prt, cx.n_tmp, cx.lna, cx.pnp, cx.pnk, cx.pos, cx.n_res,
                    cx.n_mcs, steps, dump, nc, int(exchange), seed & 0xFFFFFFFF, seed >> 32,
                    int(cx.move_kind)))
        f.write(np.asarray(cx.move_scale, np.float32).tobytes())
        f.write(np.asarray(cx.pocket_center, np.float32).tobytes())
        for a in (cx.lig_x, cx.lig_y, cx.lig_z, cx.lig_t, cx.lig_c, cx.lig_n, cx.prt_x, cx.prt_y,
                  cx.prt_z, cx.prt_t, cx.prt_c, cx.prt_ele, cx.prt_seq3r,
                  cx.prt_c0_and_d12_or_c2, cx.prt_hpp, cx.psp, cx.kde_x, cx.kde_y, cx.kde_z,
                  cx.kde_t, cx.mcs_x, cx.mcs_y, cx.mcs_z, cx.mcs_tcc):
            f.write(np.ascontiguousarray(a).tobytes())
        f.write(cx.enepara.tobytes())
        for t, mb in zip(cx.temp_t, cx.temp_minus_beta):
            f.write(np.float32(t).tobytes() + np.float32(mb).tobytes() + np.int32(0).tobytes())


@pytest.mark.gpu
def test_cpp_run_matches_c_api(tmp_path):
    import geaxdock_gpu_amd as gd
    assert os.path.exists(BIN), "build tests/cpp/run_ref_test first (__graft_entry__.build)"
    cx = synth.single_complex(seed=6, move_kind=abi.MOVE_UNIFORM, trans_sigma=0.05,
                       rot_sigma_deg=2.0, n_lig=2, n_prt=2)
    steps, dump, nc, seed = 300, 120, 20, 0x1234_5678_9ABC
    inp, out, csv = tmp_path / "in.bin", tmp_path / "out.bin", tmp_path / "states.csv"
    write_input(cx, inp, steps, dump, nc, True, seed)
    r = subprocess.run([BIN, str(inp), str(out), str(csv)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    raw = out.read_bytes()
    n = int(np.frombuffer(raw[:8], np.int64)[0])
    recs = np.frombuffer(raw[8:8 + n * abi.RECORD_DTYPE.itemsize], abi.RECORD_DTYPE)
    t0, t1, ar = np.frombuffer(raw[8 + n * abi.RECORD_DTYPE.itemsize:], np.float64)

    got = []
    with gd.Engine(seed=seed, steps_per_dump=dump) as eng:
        eng.add_complex(cx)
        eng.run(steps, nc, sink=lambda c, rr: got.append(rr))
    got = np.concatenate(got)
    order = np.lexsort((np.arange(len(got)), got["replica"][:, 0]))
    got = got[order]
    assert n == len(got)
    assert np.array_equal(recs.view(np.uint8), got.view(np.uint8))
    assert ar == pytest.approx(n / (cx.n_rep * steps), rel=1e-6)
    assert 0 < t0 <= t1
    lines = csv.read_text().splitlines()
    assert lines[0].split() == ["lig", "prt", "cms", "rmsd", "mv1", "mv2", "mv3", "mv4", "mv5",
                                "mv6", "vdw", "ele", "pmf", "psp", "hdb", "hpc", "kde", "lhm",
                                "edst", "ener"]
    first = recs[recs["replica"][:, 0] == recs["replica"][0, 0]]
    assert len(lines) == 1 + len(first)
    vals = [float(v) for v in lines[1].split()]
    assert vals[19] == pytest.approx(float(first[0]["energy"]["e"][9]), abs=1e-4)
