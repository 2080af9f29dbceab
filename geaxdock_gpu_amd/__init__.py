"""MI355X replica-exchange Monte Carlo docking engine (GeauxDock hot path).

Python mirror of the engine's C ABI (include/geauxdock.h). `run()` follows the
reference's Run() (src/run.h:10-22, src/run.cu:33-289): it takes one complex,
runs the MC launch loop and returns the accepted-step records per replica and
the McLog figures (t0 compute time, t1 wall time, acceptance ratio,
steps_total) that PrintSummary reports (src/util.C:1347-1364).

The library is the in-tree libgeauxdock.so; there is no fallback path.
"""
import ctypes as C
import os

import numpy as np

from . import abi

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgeauxdock.so")

__all__ = ["Engine", "GeauxDockError", "load_library", "run", "summary", "abi"]


class GeauxDockError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"{abi.STATUS.get(status, status)}: {msg}")
        self.status = status


_lib = None


def load_library():
    """Load the HIP engine. Raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with geaxdock_gpu_amd.build.build()")
    lib = C.CDLL(LIB_PATH)
    P = C.POINTER
    vp = C.c_void_p
    lib.gd_create.argtypes = [P(vp), P(abi.Opts)]
    lib.gd_destroy.argtypes = [vp]
    lib.gd_destroy.restype = None
    lib.gd_last_error.argtypes = [vp]
    lib.gd_last_error.restype = C.c_char_p
    lib.gd_add_complex.argtypes = [vp, P(abi.ComplexDesc), P(C.c_int)]
    lib.gd_run.argtypes = [vp, C.c_int64, C.c_int, abi.RECORD_SINK, vp, P(abi.RunStats)]
    lib.gd_get_state.argtypes = [vp, C.c_int, P(C.c_float), vp, vp, P(C.c_int32), P(C.c_int32)]
    lib.gd_checkpoint.argtypes = [vp, vp, P(C.c_size_t)]
    lib.gd_restore.argtypes = [vp, vp, C.c_size_t]
    lib.gd_calc_energy.argtypes = [vp, C.c_int, C.c_int, P(C.c_int), P(C.c_int), P(C.c_float), vp]
    lib.gd_calc_energy_noalloc.argtypes = [vp, C.c_int, C.c_int, P(C.c_int), P(C.c_int),
                                           P(C.c_float), vp]
    lib.gd_calc_energy_noalloc.restype = C.c_int
    lib.gd_accept.argtypes = [vp, C.c_int, P(C.c_float), P(C.c_float), P(C.c_float),
                              P(C.c_float), P(C.c_int32)]
    lib.gd_exchange.argtypes = [vp, C.c_int, C.c_int, C.c_int, P(C.c_float), P(C.c_float),
                                P(C.c_float), P(C.c_int32), P(C.c_int32)]
    lib.gd_philox.argtypes = [vp, C.c_int, P(C.c_uint32), P(C.c_uint32), P(C.c_uint32)]
    lib.gd_num_replicas.argtypes = [vp, C.c_int]
    lib.gd_num_replicas.restype = C.c_int
    lib.gd_get_best.argtypes = [vp, C.c_int, P(C.c_float), P(C.c_float)]
    lib.gd_get_best.restype = C.c_int
    lib.gd_reset_best.argtypes = [vp]
    lib.gd_reset_best.restype = C.c_int
    lib.gd_complex_shape.argtypes = [vp, C.c_int, P(C.c_int), P(C.c_int)]
    lib.gd_complex_shape.restype = C.c_int
    lib.gd_run_da.argtypes = [vp, C.c_int64, C.c_int, C.c_int, C.c_float, abi.RECORD_SINK, vp,
                              P(abi.RunStats)]
    lib.gd_run_da.restype = C.c_int
    lib.gd_sur_tabulate.argtypes = [vp]
    lib.gd_sur_tabulate.restype = C.c_int
    lib.gd_sur_surrogate.argtypes = [vp, C.c_int, C.c_int, P(C.c_int), P(C.c_int), P(C.c_float),
                                     C.c_float, P(C.c_float)]
    lib.gd_sur_surrogate.restype = C.c_int
    lib.gd_sur_num_params.argtypes = []
    lib.gd_sur_num_params.restype = C.c_int
    lib.gd_sur_build_grids.argtypes = [vp, C.c_float, C.c_float]
    lib.gd_sur_features.argtypes = [vp, C.c_int, C.c_int, P(C.c_int), P(C.c_int), P(C.c_float),
                                    P(C.c_float)]
    lib.gd_sur_set_params.argtypes = [vp, P(C.c_float)]
    lib.gd_sur_get_params.argtypes = [vp, P(C.c_float)]
    lib.gd_sur_dataset_clear.argtypes = [vp]
    lib.gd_sur_dataset_reserve.argtypes = [vp, C.c_int64]
    lib.gd_sur_dataset_add.argtypes = [vp, C.c_int, C.c_int, P(C.c_int), P(C.c_int),
                                       P(C.c_float), P(C.c_float)]
    lib.gd_sur_dataset_add_multi.argtypes = [vp, C.c_int, P(C.c_int), P(C.c_int), P(C.c_int),
                                             P(C.c_float), P(C.c_float)]
    lib.gd_sur_dataset_add_multi.restype = C.c_int
    lib.gd_sur_dataset_size.argtypes = [vp]
    lib.gd_sur_dataset_size.restype = C.c_int64
    lib.gd_sur_loss_grad.argtypes = [vp, C.c_int, C.c_int, P(C.c_double), P(C.c_float)]
    lib.gd_sur_predict.argtypes = [vp, C.c_int, C.c_int, P(C.c_float)]
    lib.gd_sur_train.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_float,
                                 P(C.c_double), P(C.c_double)]
    lib.gd_kmeans_medoids.argtypes = [vp, C.c_int, P(C.c_int64), P(C.c_float), C.c_int,
                                      C.c_float, C.c_int, C.c_uint64, P(C.c_int), P(C.c_int),
                                      P(C.c_int), P(C.c_double)]
    lib.gd_post_mc.argtypes = [vp, vp, C.c_int64, C.c_int, C.c_int, C.c_uint64, P(C.c_int64),
                               P(C.c_int), P(C.c_int), P(C.c_double)]
    for f in ("gd_kmeans_medoids", "gd_post_mc", "gd_create", "gd_add_complex", "gd_run", "gd_get_state", "gd_checkpoint",
              "gd_restore", "gd_calc_energy", "gd_accept", "gd_exchange", "gd_philox",
              "gd_sur_build_grids", "gd_sur_features", "gd_sur_set_params", "gd_sur_get_params",
              "gd_sur_dataset_clear", "gd_sur_dataset_reserve", "gd_sur_dataset_add",
              "gd_sur_loss_grad", "gd_sur_predict", "gd_sur_train"):
        getattr(lib, f).restype = C.c_int
    _lib = lib
    return lib


def _f(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _i(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


def _u(a):
    return a.ctypes.data_as(C.POINTER(C.c_uint32))


def _records(ptr, n):
    """Copy of n gd_record entries at ptr (the library's buffer is reused after
    the sink returns)."""
    raw = np.ctypeslib.as_array(C.cast(ptr, C.POINTER(C.c_uint8)),
                                shape=(n * abi.RECORD_DTYPE.itemsize,))
    return raw.view(abi.RECORD_DTYPE).copy()


class Engine:
    """One engine context on one device (gd_create ... gd_destroy)."""

    def __init__(self, seed=0, device=0, steps_per_dump=2000, record=True, lanes_per_replica=0,
                 atoms_per_lane=0, verbose=False, receptor_split=0, ladders_per_pack=0,
                 launch=0, receptor_tile=0):
        self.lib = load_library()
        o = abi.Opts()
        o.device, o.seed, o.steps_per_dump = device, seed, steps_per_dump
        o.record, o.verbose = int(record), int(verbose)
        o.lanes_per_replica, o.atoms_per_lane = lanes_per_replica, atoms_per_lane
        o.receptor_split = receptor_split
        o.ladders_per_pack = ladders_per_pack
        o.launch = launch
        o.receptor_tile = receptor_tile
        self._ctx = C.c_void_p()
        st = self.lib.gd_create(C.byref(self._ctx), C.byref(o))
        if st != 0:
            msg = self._err()
            self.close()
            raise GeauxDockError(st, msg)
        self._descs = []
        self.complexes = []

    def _err(self):
        if not self._ctx:
            return "no context"
        return self.lib.gd_last_error(self._ctx).decode()

    def _check(self, st):
        if st != 0:
            raise GeauxDockError(st, self._err())

    def close(self):
        if self._ctx:
            self.lib.gd_destroy(self._ctx)
            self._ctx = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_complex(self, cx):
        d = cx.desc()
        cid = C.c_int(-1)
        self._check(self.lib.gd_add_complex(self._ctx, C.byref(d), C.byref(cid)))
        self._descs.append(d)
        self.complexes.append(cx)
        return cid.value

    def run(self, n_steps, steps_per_exchange=0, sink=None):
        """Advance every replica n_steps. sink(cid, records ndarray) gets each
        dump window's accepted-state records."""
        st = abi.RunStats()
        if sink is None:
            cb = abi.RECORD_SINK()
        else:
            def _cb(user, cid, recs, n):
                sink(cid, _records(recs, n))
            cb = abi.RECORD_SINK(_cb)
        self._check(self.lib.gd_run(self._ctx, int(n_steps), int(steps_per_exchange), cb, None,
                                    C.byref(st)))
        return {f: getattr(st, f) for f, _ in abi.RunStats._fields_}

    def run_da(self, n_steps, steps_per_exchange, k_sub, sur_scale, sink=None):
        """Delayed-acceptance run: k_sub surrogate steps per exact evaluation."""
        st = abi.RunStats()
        if sink is None:
            cb = abi.RECORD_SINK()
        else:
            def _cb(user, cid, recs, n):
                sink(cid, _records(recs, n))
            cb = abi.RECORD_SINK(_cb)
        self._check(self.lib.gd_run_da(self._ctx, int(n_steps), int(steps_per_exchange),
                                       int(k_sub), float(sur_scale), cb, None, C.byref(st)))
        return {f: getattr(st, f) for f, _ in abi.RunStats._fields_}

    def best(self, cid):
        """(best total energy [n_rep], its movematrix [n_rep, 6]) per replica."""
        n = self.num_replicas(cid)
        e = np.zeros(n, np.float32)
        mm = np.zeros((n, 6), np.float32)
        self._check(self.lib.gd_get_best(self._ctx, cid, _f(e), _f(mm)))
        return e, mm

    def reset_best(self):
        self._check(self.lib.gd_reset_best(self._ctx))

    def shape(self, cid):
        """(lanes per replica, ligand atoms per lane) chosen for a complex."""
        lpr, apl = C.c_int(0), C.c_int(0)
        self._check(self.lib.gd_complex_shape(self._ctx, cid, C.byref(lpr), C.byref(apl)))
        return lpr.value, apl.value

    def num_replicas(self, cid):
        return self.lib.gd_num_replicas(self._ctx, cid)

    def state(self, cid):
        n = self.num_replicas(cid)
        mm = np.zeros((n, 6), np.float32)
        e = np.zeros(n, abi.ENERGY_DTYPE)
        rep = np.zeros((n, 4), np.int32)
        acs_mc = np.zeros(n, np.int32)
        acs_x = np.zeros(n, np.int32)
        self._check(self.lib.gd_get_state(self._ctx, cid, _f(mm), e.ctypes.data, rep.ctypes.data,
                                          _i(acs_mc), _i(acs_x)))
        return dict(movematrix=mm, energy=e, replica=rep, acs_mc=acs_mc, acs_x=acs_x)

    def checkpoint(self):
        sz = C.c_size_t(0)
        self._check(self.lib.gd_checkpoint(self._ctx, None, C.byref(sz)))
        buf = C.create_string_buffer(sz.value)
        self._check(self.lib.gd_checkpoint(self._ctx, buf, C.byref(sz)))
        return bytes(buf.raw[: sz.value])

    def restore(self, blob):
        buf = C.create_string_buffer(blob, len(blob))
        self._check(self.lib.gd_restore(self._ctx, buf, len(blob)))

    def calc_energy(self, cid, idx_lig, idx_prt, movematrix):
        idx_lig = np.ascontiguousarray(idx_lig, np.int32)
        idx_prt = np.ascontiguousarray(idx_prt, np.int32)
        mm = np.ascontiguousarray(movematrix, np.float32).reshape(-1, 6)
        n = mm.shape[0]
        out = np.zeros(n, abi.ENERGY_DTYPE)
        self._check(self.lib.gd_calc_energy(self._ctx, cid, n, _i(idx_lig), _i(idx_prt), _f(mm),
                                            out.ctypes.data))
        return out

    def calc_energy_noalloc(self, cid, idx_lig, idx_prt, movematrix):
        """calc_energy through gd_calc_energy_noalloc: no device allocation,
        at most abi.NOALLOC_MAX_POSES poses; needs one calc_energy or run
        call after the last add_complex."""
        idx_lig = np.ascontiguousarray(idx_lig, np.int32)
        idx_prt = np.ascontiguousarray(idx_prt, np.int32)
        mm = np.ascontiguousarray(movematrix, np.float32).reshape(-1, 6)
        n = mm.shape[0]
        out = np.zeros(n, abi.ENERGY_DTYPE)
        self._check(self.lib.gd_calc_energy_noalloc(self._ctx, cid, n, _i(idx_lig), _i(idx_prt),
                                                    _f(mm), out.ctypes.data))
        return out

    def accept(self, e_new, e_old, minus_beta, u):
        a = [np.ascontiguousarray(v, np.float32) for v in (e_new, e_old, minus_beta, u)]
        out = np.zeros(len(a[0]), np.int32)
        self._check(self.lib.gd_accept(self._ctx, len(a[0]), _f(a[0]), _f(a[1]), _f(a[2]),
                                       _f(a[3]), _i(out)))
        return out

    def exchange(self, n_tmp, mode_t, etot, minus_beta, u, idx_tmp, acs):
        etot = np.ascontiguousarray(etot, np.float32)
        mb = np.ascontiguousarray(minus_beta, np.float32)
        u = np.ascontiguousarray(u, np.float32)
        it = np.array(idx_tmp, np.int32, copy=True)
        ac = np.array(acs, np.int32, copy=True)
        n_groups = it.size // n_tmp
        self._check(self.lib.gd_exchange(self._ctx, n_groups, n_tmp, mode_t, _f(etot), _f(mb),
                                         _f(u), _i(it), _i(ac)))
        return it, ac

    # ---- learned surrogate -------------------------------------------------
    @property
    def sur_num_params(self):
        return self.lib.gd_sur_num_params()

    def sur_build_grids(self, spacing=0.5, half_width=16.0):
        self._check(self.lib.gd_sur_build_grids(self._ctx, spacing, half_width))

    def sur_features(self, cid, idx_lig, idx_prt, movematrix):
        il = np.ascontiguousarray(idx_lig, np.int32)
        ip = np.ascontiguousarray(idx_prt, np.int32)
        mm = np.ascontiguousarray(movematrix, np.float32).reshape(-1, 6)
        lna = self.complexes[cid].lna
        out = np.zeros((len(mm), lna, 56), np.float32)
        self._check(self.lib.gd_sur_features(self._ctx, cid, len(mm), _i(il), _i(ip), _f(mm),
                                             _f(out)))
        return out

    def sur_tabulate(self):
        """Tabulate the network per atom type on the surrogate grids."""
        self._check(self.lib.gd_sur_tabulate(self._ctx))

    def sur_surrogate(self, cid, idx_lig, idx_prt, movematrix, sur_scale):
        """Surrogate totals of poses as the delayed-acceptance kernel sees them."""
        il = np.ascontiguousarray(idx_lig, np.int32)
        ip = np.ascontiguousarray(idx_prt, np.int32)
        mm = np.ascontiguousarray(movematrix, np.float32).reshape(-1, 6)
        out = np.zeros(len(mm), np.float32)
        self._check(self.lib.gd_sur_surrogate(self._ctx, cid, len(mm), _i(il), _i(ip), _f(mm),
                                              float(sur_scale), _f(out)))
        return out

    def sur_set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == self.sur_num_params
        self._check(self.lib.gd_sur_set_params(self._ctx, _f(p)))

    def sur_get_params(self):
        p = np.zeros(self.sur_num_params, np.float32)
        self._check(self.lib.gd_sur_get_params(self._ctx, _f(p)))
        return p

    def sur_dataset_clear(self):
        self._check(self.lib.gd_sur_dataset_clear(self._ctx))

    def sur_dataset_reserve(self, rows):
        self._check(self.lib.gd_sur_dataset_reserve(self._ctx, int(rows)))

    def sur_dataset_add(self, cid, idx_lig, idx_prt, movematrix, target):
        il = np.ascontiguousarray(idx_lig, np.int32)
        ip = np.ascontiguousarray(idx_prt, np.int32)
        mm = np.ascontiguousarray(movematrix, np.float32).reshape(-1, 6)
        t = np.ascontiguousarray(target, np.float32)
        self._check(self.lib.gd_sur_dataset_add(self._ctx, cid, len(mm), _i(il), _i(ip), _f(mm),
                                                _f(t)))

    def sur_dataset_add_multi(self, cids, idx_lig, idx_prt, movematrix, target):
        """Append poses of several complexes in the given order."""
        cc = np.ascontiguousarray(cids, np.int32)
        il = np.ascontiguousarray(idx_lig, np.int32)
        ip = np.ascontiguousarray(idx_prt, np.int32)
        mm = np.ascontiguousarray(movematrix, np.float32).reshape(-1, 6)
        t = np.ascontiguousarray(target, np.float32)
        self._check(self.lib.gd_sur_dataset_add_multi(self._ctx, len(mm), _i(cc), _i(il), _i(ip),
                                                      _f(mm), _f(t)))

    def sur_dataset_size(self):
        return self.lib.gd_sur_dataset_size(self._ctx)

    def sur_loss_grad(self, begin, end, grad=True):
        loss = C.c_double(0.0)
        g = np.zeros(self.sur_num_params, np.float32) if grad else None
        self._check(self.lib.gd_sur_loss_grad(self._ctx, begin, end, C.byref(loss),
                                              _f(g) if grad else None))
        return loss.value, g

    def sur_predict(self, begin, end):
        out = np.zeros(end - begin, np.float32)
        self._check(self.lib.gd_sur_predict(self._ctx, begin, end, _f(out)))
        return out

    def sur_train(self, begin, end, epochs, batch_poses, lr=1e-3):
        hist = np.zeros(max(epochs, 1), np.float64)
        secs = C.c_double(0.0)
        self._check(self.lib.gd_sur_train(self._ctx, begin, end, epochs, batch_poses, lr,
                                          hist.ctypes.data_as(C.POINTER(C.c_double)),
                                          C.byref(secs)))
        return hist[:epochs], secs.value

    def philox(self, ctr, key):
        ctr = np.ascontiguousarray(ctr, np.uint32).reshape(-1, 4)
        key = np.ascontiguousarray(key, np.uint32).reshape(-1, 2)
        out = np.zeros_like(ctr)
        self._check(self.lib.gd_philox(self._ctx, ctr.shape[0], _u(ctr), _u(key), _u(out)))
        return out

    def kmeans_medoids(self, features, group_off, k, threshold=0.001, max_pass=501, seed=0,
                       timing=None):
        """Device k-means per group (gd_kmeans_medoids): (medoid[g][k] group-local
        index or -1, size[g][k], passes[g]). `timing` (a dict) receives the
        device milliseconds."""
        f = np.ascontiguousarray(features, np.float32).reshape(-1, abi.MAXWEI - 1)
        off = np.ascontiguousarray(group_off, np.int64)
        g = len(off) - 1
        med = np.zeros((max(g, 1), k), np.int32)
        sz = np.zeros((max(g, 1), k), np.int32)
        ps = np.zeros(max(g, 1), np.int32)
        ms = C.c_double(0.0)
        self._check(self.lib.gd_kmeans_medoids(
            self._ctx, g, off.ctypes.data_as(C.POINTER(C.c_int64)), _f(f), int(k),
            float(threshold), int(max_pass), int(seed), _i(med), _i(sz), _i(ps), C.byref(ms)))
        if timing is not None:
            timing["device_ms"] = ms.value
        return med[:g], sz[:g], ps[:g]

    def post_mc(self, records, k_replica=50, k_final=500, seed=0, timing=None):
        """post_mc on the device (gd_post_mc): (record indices of the medoids,
        their cluster sizes). `timing` (a dict) receives the device
        milliseconds of both levels."""
        r = np.ascontiguousarray(records, abi.RECORD_DTYPE)
        mr = np.zeros(k_final, np.int64)
        ms = np.zeros(k_final, np.int32)
        m = C.c_int(0)
        dms = C.c_double(0.0)
        self._check(self.lib.gd_post_mc(self._ctx, r.ctypes.data, len(r), int(k_replica),
                                        int(k_final), int(seed),
                                        mr.ctypes.data_as(C.POINTER(C.c_int64)), _i(ms),
                                        C.byref(m), C.byref(dms)))
        if timing is not None:
            timing["device_ms"] = dms.value
        return mr[:m.value], ms[:m.value]


def run(cx, steps_total, steps_per_exchange=20, seed=0, steps_per_dump=2000, device=0):
    """Run() for one complex: returns (records by replica, mclog).

    mclog holds t0 (MC-kernel seconds), t1 (wall seconds), ar (accepted
    records / (n_rep * steps_total)) and steps_total, as McLog does.
    """
    records = {}

    def sink(cid, recs):
        for r in np.unique(recs["replica"][:, 0]):
            records.setdefault(int(r), []).append(recs[recs["replica"][:, 0] == r])

    with Engine(seed=seed, device=device, steps_per_dump=steps_per_dump) as eng:
        eng.add_complex(cx)
        st = eng.run(steps_total, steps_per_exchange, sink=sink)
    recs = {k: np.concatenate(v) for k, v in records.items()}
    n_rec = sum(len(v) for v in recs.values())
    mclog = dict(t0=st["t_compute"], t1=st["t_wall"], steps_total=steps_total,
                 ar=n_rec / float(cx.n_rep * max(steps_total, 1)),
                 mc_sweeps_per_s=steps_total * cx.n_rep / st["t_compute"] if st["t_compute"] else 0.0)
    return recs, mclog


def summary(mclog, cx):
    """The performance part of PrintSummary (src/util.C:1281-1364) as text:
    sizes, the acceptance ratio, and replica-steps ("MC sweeps") per second
    over compute time (t0) and wall time (t1)."""
    bar = "=" * 80
    n_rep = cx.n_lig * cx.n_prt * cx.n_tmp
    lines = [bar, f"lig conformations\t\t{cx.n_lig}", f"prt conformations\t\t{cx.n_prt}",
             f"temperatures\t\t\t{cx.n_tmp}", f"replica ensembles\t\t{n_rep}",
             f"size_lig\t\t\t{cx.lna}", f"size_prt\t\t\t{cx.pnp}", f"size_pnk\t\t\t{cx.pnk}",
             f"size_mcs\t\t\t{cx.pos}", f"AR of MC\t\t\t{mclog['ar']:.3f}", bar, "Performance", bar]
    n = float(mclog["steps_total"]) * n_rep
    for name, t in (("compute time", mclog["t0"]), ("wall time", mclog["t1"])):
        rate = n / t if t > 0 else 0.0
        lines += [f"{name}\t\t\t{t:.3f} seconds",
                  f"time per MC sweep per replica\t{(1e6 / rate if rate else 0.0):.3f} * 1e-6 seconds",
                  f"MC sweeps per second\t\t{rate:.3f}", f"speedup over 843.75\t\t{rate / 843.75:.3f} X"]
    lines.append(bar)
    return "\n".join(lines)
