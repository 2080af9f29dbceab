"""The reference's `dock` pipeline as one call (src/dock.C:21-220).

load inputs -> Optimize* / SetTemperature -> Run() -> post_mc -> printStates
-> PrintSummary, each step on this library: the loaders and the writers are
host code in libgeauxdock.so, the Monte Carlo run and the k-means medoid
selection run on the GPU. The command line itself (ParseArguments) is not
reproduced; the keyword arguments carry its options.
"""
import numpy as np

from . import Engine, abi, io, summary


def dock(ligand_sdf, protein_pdb, lhm_ff, paras, complex_id="", steps_total=3000,
         steps_per_exchange=10, n_tmp=1, floor_temp=0.044, ceiling_temp=0.036, ts=0.02, rs=0.08,
         move_kind=abi.MOVE_UNIFORM, seed=0, weight=None, nor_a=None, nor_b=None,
         csv_path=None, records_path=None, k_replica=50, k_final=500, steps_per_dump=2000,
         print_summary=False):
    """Dock one complex. Defaults follow the reference's data/dock.bash
    (10 steps per exchange, one temperature, move scales 0.02 / 0.08).

    Returns a dict: the complex, every accepted-step record, the medoid
    records with their cluster sizes, McLog figures and the summary text.
    csv_path: printStates of the medoids (the reference writes it to the
    trace path); records_path: every record in the flat record file."""
    cx = io.load_complex(ligand_sdf, protein_pdb, lhm_ff, paras, complex_id=complex_id,
                         n_tmp=n_tmp, floor_temp=floor_temp, ceiling_temp=ceiling_temp, ts=ts,
                         rs=rs, move_kind=move_kind, weight=weight, nor_a=nor_a, nor_b=nor_b)
    chunks = []
    rf = io.RecordFile(records_path) if records_path else None

    def sink(cid, recs):
        chunks.append(np.array(recs, copy=True))
        if rf is not None:
            rf.write(recs)

    with Engine(seed=seed, steps_per_dump=steps_per_dump) as eng:
        eng.add_complex(cx)
        st = eng.run(steps_total, steps_per_exchange, sink=sink)
        records = np.concatenate(chunks) if chunks else np.zeros(0, abi.RECORD_DTYPE)
        idx, sizes = eng.post_mc(records, k_replica, k_final, seed=seed) if len(records) else \
            (np.zeros(0, np.int64), np.zeros(0, np.int32))
    medoids = records[idx]
    if csv_path:
        io.write_states_csv(csv_path, medoids)
    mclog = dict(t0=st["t_compute"], t1=st["t_wall"], steps_total=steps_total,
                 ar=len(records) / float(cx.n_rep * max(steps_total, 1)))
    text = summary(mclog, cx)
    if print_summary:
        print(text)
    return dict(complex=cx, records=records, medoids=medoids, cluster_size=sizes, mclog=mclog,
                summary=text, stats=st)
