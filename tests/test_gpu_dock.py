"""The whole `dock` pipeline on a real complex (the reference's data/dock.bash
inputs for 1a07C1): loaders -> MC on the GPU -> records file -> post_mc on the
GPU -> printStates CSV -> PrintSummary."""
import os

import numpy as np
import pytest

from geaxdock_gpu_amd import io

pytestmark = pytest.mark.gpu

G = os.path.join(os.path.dirname(__file__), "golden", "inputs")


def test_dock_1a07C1_end_to_end(tmp_path):
    import geaxdock_gpu_amd as gd
    from geaxdock_gpu_amd.dock import dock
    csv = str(tmp_path / "1a07C1.csv")
    rec = str(tmp_path / "1a07C1.gdrec")
    out = dock(os.path.join(G, "1a07C1", "1a07C1.sdf"), os.path.join(G, "1a07C1", "1a07C.pdb"),
               os.path.join(G, "1a07C1", "1a07C1-0.8.ff"), os.path.join(G, "parameters", "paras"),
               complex_id="1a07C1", steps_total=2000, steps_per_exchange=10, n_tmp=2,
               floor_temp=0.044, ceiling_temp=0.36, seed=5, csv_path=csv, records_path=rec)
    cx, records, medoids = out["complex"], out["records"], out["medoids"]
    assert cx.n_rep == 21 * 3 * 2
    # every replica records its start pose and its accepted steps
    assert len(records) >= cx.n_rep and set(records["replica"][:, 0]) == set(range(cx.n_rep))
    assert len(io.RecordFile(rec, append=True)) == len(records)
    assert io.RecordFile(rec, append=True).read().tobytes() == records.tobytes()
    # medoids: at most 500, sizes count the level-1 medoids
    assert 0 < len(medoids) <= 500
    assert out["cluster_size"].sum() <= cx.n_rep * 50
    lines = open(csv).read().strip().split("\n")
    assert len(lines) == len(medoids) + 1
    got = np.array([[float(v) for v in l.split()[10:20]] for l in lines[1:]])
    np.testing.assert_allclose(got, medoids["energy"]["e"], atol=5.1e-5)
    # recorded energies are the exact energies of the recorded poses
    with gd.Engine(seed=0) as eng:
        cid = eng.add_complex(cx)
        sample = records[:: max(1, len(records) // 400)]
        again = eng.calc_energy(cid, sample["replica"][:, 3].copy(), sample["replica"][:, 1].copy(),
                                sample["movematrix"].copy())
    np.testing.assert_allclose(again["e"], sample["energy"]["e"], rtol=1e-5, atol=1e-5)
    assert "MC sweeps per second" in out["summary"]
    print("\n" + out["summary"])
