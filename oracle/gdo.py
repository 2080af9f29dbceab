"""ctypes binding of the CPU oracle (oracle/gdo.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke check and the
benchmark's cpu_baseline leg, never by the engine.
"""
import ctypes as C
import os
import subprocess

import numpy as np

from geaxdock_gpu_amd import abi

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def load():
    if not os.path.exists(_LIB):
        build()
    lib = C.CDLL(_LIB)
    u32p = C.POINTER(C.c_uint32)
    fp = C.POINTER(C.c_float)
    lib.gdo_philox4x32_10.argtypes = [u32p, u32p, u32p]
    lib.gdo_uniform.argtypes = [C.c_uint32]
    lib.gdo_uniform.restype = C.c_float
    lib.gdo_mc_uniforms.argtypes = [C.c_uint64, C.c_int, C.c_int, C.c_int, fp]
    lib.gdo_exchange_uniform.argtypes = [C.c_uint64, C.c_int, C.c_int, C.c_int, C.c_int]
    lib.gdo_exchange_uniform.restype = C.c_float
    lib.gdo_move_offsets.argtypes = [fp, fp, C.c_int, fp]
    lib.gdo_expf.argtypes = [C.c_float]
    lib.gdo_expf.restype = C.c_float
    lib.gdo_accept.argtypes = [C.c_float, C.c_float, C.c_float, C.c_float]
    lib.gdo_accept.restype = C.c_int
    lib.gdo_exchange_group.argtypes = [C.c_int, C.c_int, fp, fp, fp,
                                       C.POINTER(C.c_int32), C.POINTER(C.c_int32)]
    lib.gdo_refmap_f64.argtypes = [C.POINTER(abi.ComplexDesc), C.c_void_p]
    lib.gdo_refmap_f32.argtypes = [C.POINTER(abi.ComplexDesc), C.c_void_p]
    lib.gdo_energy_f64.argtypes = [C.POINTER(abi.ComplexDesc), C.c_void_p, C.c_int, C.c_int,
                                   fp, C.POINTER(Energy64)]
    lib.gdo_energy_f32.argtypes = [C.POINTER(abi.ComplexDesc), C.c_void_p, C.c_int, C.c_int,
                                   fp, C.POINTER(abi.Energy)]
    lib.gdo_energy_f64_batch.argtypes = [C.POINTER(abi.ComplexDesc), C.c_void_p, C.c_int,
                                         C.POINTER(C.c_int), C.POINTER(C.c_int), fp, C.c_void_p,
                                         C.c_int]
    lib.gdo_place_f64.argtypes = [C.POINTER(abi.ComplexDesc), C.c_int, fp,
                                  C.POINTER(C.c_double), C.POINTER(C.c_double),
                                  C.POINTER(C.c_double)]
    lib.gdo_mc_run.argtypes = [C.POINTER(abi.ComplexDesc), C.c_uint64, C.c_int, C.c_int,
                               C.c_int, C.c_int, C.c_int, C.POINTER(McIO), C.c_int]
    lib.gdo_mc_run.restype = C.c_int
    lib.gdo_mc_run_da.argtypes = [C.POINTER(abi.ComplexDesc), C.c_uint64, C.c_int, C.c_int,
                                  C.c_int, C.c_int, C.c_int, C.POINTER(McIO), C.c_int]
    lib.gdo_mc_run_da.restype = C.c_int
    ip = C.POINTER(C.c_int)
    lib.gdo_km_init.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_uint64, ip]
    lib.gdo_kmeans.argtypes = [fp, C.c_int, C.c_int, C.c_float, C.c_int, C.c_int, C.c_int,
                               C.c_uint64, ip, ip, ip]
    lib.gdo_kmeans.restype = C.c_int
    lib.gdo_post_mc.argtypes = [C.c_void_p, C.c_int64, C.c_int, C.c_int, C.c_uint64,
                                C.POINTER(C.c_int64), ip]
    lib.gdo_post_mc.restype = C.c_int
    return lib


class Energy64(C.Structure):
    _fields_ = [("e", C.c_double * 10), ("raw", C.c_double * 10), ("mag", C.c_double * 10),
                ("cms", C.c_double), ("rmsd", C.c_double), ("tp", C.c_int64),
                ("fn", C.c_int64), ("fp", C.c_int64), ("tn", C.c_int64),
                ("min_margin", C.c_double)]


class McIO(C.Structure):
    _fields_ = [("mm", C.POINTER(C.c_float)), ("e_old", C.POINTER(abi.Energy)),
                ("idx_tmp", C.POINTER(C.c_int32)), ("acs_mc", C.POINTER(C.c_int32)),
                ("acs_x", C.POINTER(C.c_int32)), ("rec", C.POINTER(abi.Record)),
                ("rec_cap", C.c_int64), ("rec_cnt", C.POINTER(C.c_int64)),
                ("mlog", C.POINTER(C.c_float)), ("clog", C.POINTER(C.c_float)),
                ("xlog", C.POINTER(C.c_float)), ("log_step0", C.c_int),
                ("log_steps", C.c_int), ("log_segs", C.c_int)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        _lib = load()
    return _lib


def philox(ctr, key):
    ctr = np.ascontiguousarray(ctr, np.uint32)
    key = np.ascontiguousarray(key, np.uint32)
    out = np.zeros(4, np.uint32)
    u32p = C.POINTER(C.c_uint32)
    lib().gdo_philox4x32_10(ctr.ctypes.data_as(u32p), key.ctypes.data_as(u32p),
                            out.ctypes.data_as(u32p))
    return out


def expf(x):
    return np.float32(lib().gdo_expf(float(np.float32(x))))


def refmap_f64(desc, cx):
    ref = np.zeros(cx.lna * cx.pnp, np.uint8)
    lib().gdo_refmap_f64(C.byref(desc), ref.ctypes.data)
    return ref


def refmap_f32(desc, cx):
    ref = np.zeros(cx.lna * cx.pnp, np.uint8)
    lib().gdo_refmap_f32(C.byref(desc), ref.ctypes.data)
    return ref


def energy_f64(desc, ref, iprt, ilig, mm):
    out = Energy64()
    mm = np.ascontiguousarray(mm, np.float32)
    lib().gdo_energy_f64(C.byref(desc), None if ref is None else ref.ctypes.data, iprt, ilig,
                         abi.fptr(mm), C.byref(out))
    return out


ENERGY64_DTYPE = np.dtype([("e", "<f8", 10), ("raw", "<f8", 10), ("mag", "<f8", 10),
                           ("cms", "<f8"), ("rmsd", "<f8"), ("tp", "<i8"), ("fn", "<i8"),
                           ("fp", "<i8"), ("tn", "<i8"), ("min_margin", "<f8")])


def energy_f64_batch(desc, ref, iprt, ilig, mm, threads=0):
    """energy_f64 of many poses (OpenMP over poses), as an ENERGY64_DTYPE array."""
    iprt = np.ascontiguousarray(iprt, np.int32)
    ilig = np.ascontiguousarray(ilig, np.int32)
    mm = np.ascontiguousarray(mm, np.float32).reshape(-1, 6)
    out = np.zeros(len(mm), ENERGY64_DTYPE)
    assert ENERGY64_DTYPE.itemsize == C.sizeof(Energy64)
    ip = C.POINTER(C.c_int)
    lib().gdo_energy_f64_batch(C.byref(desc), None if ref is None else ref.ctypes.data, len(mm),
                               iprt.ctypes.data_as(ip), ilig.ctypes.data_as(ip), abi.fptr(mm),
                               out.ctypes.data, threads)
    return out


def energy_f32(desc, ref, iprt, ilig, mm):
    out = abi.Energy()
    mm = np.ascontiguousarray(mm, np.float32)
    lib().gdo_energy_f32(C.byref(desc), None if ref is None else ref.ctypes.data, iprt, ilig,
                         abi.fptr(mm), C.byref(out))
    return out


def place_f64(desc, cx, ilig, mm):
    mm = np.ascontiguousarray(mm, np.float32)
    xyz = np.zeros((3, cx.lna), np.float64)
    dp = C.POINTER(C.c_double)
    lib().gdo_place_f64(C.byref(desc), ilig, abi.fptr(mm), xyz[0].ctypes.data_as(dp),
                        xyz[1].ctypes.data_as(dp), xyz[2].ctypes.data_as(dp))
    return xyz


class McState:
    """Per-replica state of the CPU MC path for one complex."""

    def __init__(self, cx, rec_cap=0):
        n = cx.n_rep
        self.mm = np.zeros((n, 6), np.float32)
        self.e_old = np.zeros(n, abi.ENERGY_DTYPE)
        flat = np.arange(n)
        self.idx_tmp = ((flat // cx.n_lig) % cx.n_tmp).astype(np.int32)
        self.acs_mc = np.zeros(n, np.int32)
        self.acs_x = np.zeros(n, np.int32)
        self.rec_cap = rec_cap
        self.rec = np.zeros(max(1, n * rec_cap), abi.RECORD_DTYPE)
        self.rec_cnt = np.zeros(n, np.int64)
        self.n_ladders = cx.n_prt * cx.n_lig
        self.mlog = self.clog = self.xlog = None

    def enable_log(self, step0, n_steps, n_segs):
        """Log the decision margins (FP64 of the FP32 operands, positive
        accepts) of steps [step0, step0 + n_steps) and of the exchanges after
        segments 0 .. n_segs - 1: mlog[replica, step - step0] (Metropolis, or
        a delayed-acceptance sub-step), clog[replica, step - step0] (composite
        test of the DA group ending at that step) and xlog[seg, ladder, pair]."""
        n = len(self.mm)
        self.log_step0, self.log_steps, self.log_segs = int(step0), int(n_steps), int(n_segs)
        self.mlog = np.full((n, max(1, n_steps)), np.nan, np.float32)
        self.clog = np.full((n, max(1, n_steps)), np.nan, np.float32)
        self.xlog = np.full((max(1, n_segs), self.n_ladders, abi.MAX_NTMP // 2), np.nan,
                            np.float32)

    def io(self):
        io = McIO()
        io.mm = abi.fptr(self.mm)
        io.e_old = C.cast(self.e_old.ctypes.data, C.POINTER(abi.Energy))
        io.idx_tmp = self.idx_tmp.ctypes.data_as(C.POINTER(C.c_int32))
        io.acs_mc = self.acs_mc.ctypes.data_as(C.POINTER(C.c_int32))
        io.acs_x = self.acs_x.ctypes.data_as(C.POINTER(C.c_int32))
        io.rec = C.cast(self.rec.ctypes.data, C.POINTER(abi.Record)) if self.rec_cap else None
        io.rec_cap = self.rec_cap
        io.rec_cnt = self.rec_cnt.ctypes.data_as(C.POINTER(C.c_int64))
        if self.mlog is not None:
            io.mlog = abi.fptr(self.mlog)
            io.clog = abi.fptr(self.clog)
            io.xlog = abi.fptr(self.xlog)
            io.log_step0, io.log_steps, io.log_segs = (self.log_step0, self.log_steps,
                                                       self.log_segs)
        return io

    def records(self, flat):
        return self.rec[flat * self.rec_cap: flat * self.rec_cap + self.rec_cnt[flat]]


def mc_run(desc, state, seed, cid=0, init=False, step0=0, n_steps=0, nc=0, threads=0):
    io = state.io()
    rc = lib().gdo_mc_run(C.byref(desc), seed, cid, int(init), step0, n_steps, nc, C.byref(io),
                          threads)
    if rc != 0:
        raise RuntimeError(f"gdo_mc_run failed: {rc}")


def mc_run_da(desc, state, seed, cid=0, step0=0, n_steps=0, nc=0, k_sub=1, threads=0):
    """Delayed-acceptance CPU path with the surrogate reduced to its exact
    cheap terms (surrogate scale 0); state must hold evaluated start poses."""
    io = state.io()
    rc = lib().gdo_mc_run_da(C.byref(desc), seed, cid, step0, n_steps, nc, k_sub, C.byref(io),
                             threads)
    if rc != 0:
        raise RuntimeError(f"gdo_mc_run_da failed: {rc}")


def kmeans(features, k, threshold=0.001, max_pass=501, group=0, level=0, seed=0):
    """One group: (membership, medoid[k], size[k], passes)."""
    f = np.ascontiguousarray(features, np.float32).reshape(-1, 9)
    n = len(f)
    memb = np.zeros(max(n, 1), np.int32)
    med = np.zeros(k, np.int32)
    sz = np.zeros(k, np.int32)
    ip = C.POINTER(C.c_int)
    p = lib().gdo_kmeans(f.ctypes.data_as(C.POINTER(C.c_float)), n, k, threshold, max_pass,
                         group, level, seed, memb.ctypes.data_as(ip), med.ctypes.data_as(ip),
                         sz.ctypes.data_as(ip))
    return memb[:n], med, sz, p


def km_init(n, k, group=0, level=0, seed=0):
    out = np.zeros(k, np.int32)
    lib().gdo_km_init(n, k, group, level, seed, out.ctypes.data_as(C.POINTER(C.c_int)))
    return out


def post_mc(records, k_rep=50, k_final=500, seed=0):
    """(record indices of the medoids, their cluster sizes)."""
    r = np.ascontiguousarray(records)
    mr = np.zeros(k_final, np.int64)
    ms = np.zeros(k_final, np.int32)
    m = lib().gdo_post_mc(r.ctypes.data, len(r), k_rep, k_final, seed,
                          mr.ctypes.data_as(C.POINTER(C.c_int64)),
                          ms.ctypes.data_as(C.POINTER(C.c_int)))
    return mr[:m], ms[:m]
