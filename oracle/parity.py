"""Per-term energy parity of engine outputs against the FP64 oracle.

TEST INFRASTRUCTURE ONLY: used by tests/ and the smoke check.

Bound, per term i of pose p (the engine's FP32 value g, the oracle's FP64
value r, the term's magnitude m = |a| sum |pair contribution| + |b|):
    |g - r| <= 1e-4 m                      for every term,
    |g - r| <= 1e-4 |r|                    when m <= 10 |r| (no heavy cancellation),
    |rmsd_g - rmsd_r| <= 1e-4 max(1, rmsd_r),
    |cms_g - cms_r| <= 1e-5                when no pair lies within 1e-3 A of its
                                           contact distance (a contact flip there
                                           is a rounding decision).
"""
import numpy as np

TERMS = ["vdw", "ele", "pmf", "psp", "hdb", "hpc", "kde", "lhm", "dst", "total"]


def check(gpu, ref, tol=1e-4, label=""):
    """gpu: ENERGY_DTYPE array, ref: oracle ENERGY64_DTYPE array (same poses).
    Returns {term: (worst relative error, worst error / magnitude)}; raises
    AssertionError naming the first pose and term outside the bound."""
    g = gpu["e"].astype(np.float64)
    r = ref["e"]
    mag = np.maximum(ref["mag"], 1e-30)
    err = np.abs(g - r)
    cond = err / mag
    bad = np.argwhere(~(cond <= tol))
    assert len(bad) == 0, (label, "magnitude bound", TERMS[bad[0][1]], int(bad[0][0]),
                           float(g[tuple(bad[0])]), float(r[tuple(bad[0])]),
                           float(mag[tuple(bad[0])]))
    well = (ref["mag"] <= 10.0 * np.abs(r)) & (np.abs(r) > 0)
    rel = np.where(well, err / np.where(well, np.abs(r), 1.0), 0.0)
    bad = np.argwhere(~(rel <= tol))
    assert len(bad) == 0, (label, "relative bound", TERMS[bad[0][1]], int(bad[0][0]),
                           float(g[tuple(bad[0])]), float(r[tuple(bad[0])]))
    drms = np.abs(gpu["rmsd"].astype(np.float64) - ref["rmsd"])
    assert np.all(drms <= tol * np.maximum(1.0, ref["rmsd"])), (label, "rmsd", float(drms.max()))
    sharp = ref["min_margin"] > 1e-3
    dc = np.abs(gpu["cms"].astype(np.float64) - ref["cms"])[sharp]
    assert np.all(dc <= 1e-5), (label, "cms", float(dc.max()) if len(dc) else 0.0)
    return {TERMS[t]: (float(rel[:, t].max()), float(cond[:, t].max())) for t in range(10)}
