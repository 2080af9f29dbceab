"""The C ABI library: builds, loads, exports every declared entry point, and
fails cleanly (status code, message) on bad input and without a device."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import geaxdock_gpu_amd as gd
from geaxdock_gpu_amd import abi, synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    src = "".join(open(os.path.join(ROOT, "include", h)).read()
                  for h in sorted(os.listdir(os.path.join(ROOT, "include"))) if h.endswith(".h"))
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(gd_[a-z_0-9]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    lib = gd.load_library()
    names = declared_symbols()
    assert "gd_run" in names and "gd_calc_energy" in names and len(names) >= 13
    assert "gd_load_complex" in names and "gd_sdf_read" in names
    for n in names:
        assert hasattr(lib, n), n


def test_struct_layouts_match_header():
    # sizes implied by the header's field lists (all 4-byte fields)
    assert C.sizeof(abi.EnePara) == 4 * (6 * 24 * 30 + 4 + 50 + 2 + 20 + 3 * 24 + 2 + 30)
    assert C.sizeof(abi.Energy) == 48
    assert C.sizeof(abi.Record) == 4 * (4 + 12 + 6 + 1)
    assert C.sizeof(abi.Temp) == 12 and C.sizeof(abi.Replica) == 16


def _have_gpu():
    try:
        with gd.Engine():
            return True
    except gd.GeauxDockError:
        return False


def test_no_device_or_bad_input_is_an_error_not_a_crash():
    if _have_gpu():
        with gd.Engine() as eng:
            cx = synth.single_complex(seed=0)
            cx.lig_t = cx.lig_t.copy()
            cx.lig_t[3] = 99  # type out of range
            with pytest.raises(gd.GeauxDockError) as ei:
                eng.add_complex(cx)
            assert ei.value.status == 1 and "type" in str(ei.value)
    else:
        with pytest.raises(gd.GeauxDockError) as ei:
            gd.Engine()
        assert ei.value.status == 3


def test_synthetic_workloads_are_seeded():
    a, b = synth.single_complex(seed=4), synth.single_complex(seed=4)
    for k in ("prt_x", "lig_x", "prt_t", "kde_x", "mcs_x", "psp", "enepara"):
        assert np.array_equal(getattr(a, k), getattr(b, k)), k
    c = synth.single_complex(seed=5)
    assert not np.array_equal(a.prt_x, c.prt_x)
    cx = synth.ensemble(seed=1)
    assert cx.n_rep == 3520 and cx.lna == 48 and cx.pnp == 3000 and cx.pnk == 200 and cx.pos == 8
    # receptor points keep out of the 10 A void; conformers differ by ~0.5 A
    x = np.stack([cx.prt_x, cx.prt_y, cx.prt_z]).reshape(3, cx.n_prt, cx.pnp)
    r0 = np.linalg.norm(x[:, 0] - cx.pocket_center[:, None], axis=0)
    assert r0.min() >= 10.0
    assert 0.3 < np.std(x[:, 1] - x[:, 0]) < 0.7
    # ligand bonds are 1.5 A
    lig = np.stack([cx.lig_x, cx.lig_y, cx.lig_z]).reshape(3, cx.n_lig, cx.lna)
    bonds = np.linalg.norm(np.diff(lig[:, 0], axis=1), axis=0)
    assert np.allclose(bonds, 1.5, atol=1e-4)


def test_batch_workload_sizes():
    cxs = synth.uneven_batch(6, seed=2)
    for c in cxs:
        assert 12 <= c.lna <= 64 and 800 <= c.pnp <= 4000 and c.n_rep == 120


def test_summary_text_matches_reference_layout():
    """PrintSummary's performance block (src/util.C:1347-1364): replica-steps
    per second over compute and wall time and the speed-up over 843.75."""
    text = gd.summary(dict(t0=1.5, t1=2.0, ar=0.42, steps_total=5000), synth.single_complex())
    lines = text.split("\n")
    assert "replica ensembles\t\t8" in lines and "AR of MC\t\t\t0.420" in lines
    assert "MC sweeps per second\t\t26666.667" in lines  # 5000 * 8 / 1.5
    assert "speedup over 843.75\t\t31.605 X" in lines
    assert "MC sweeps per second\t\t20000.000" in lines  # wall time
