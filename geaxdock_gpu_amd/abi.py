"""ctypes mirror of include/geauxdock.h.

The structures keep the reference's field names and order (EnePara,
src/dock_soa.h:99-133; Temp, src/dock.h:288-293; Replica, src/dock.h:302-309;
Energy, src/dock.h:108-128; LigRecordSingleStep, src/dock.h:361-367).
"""
import ctypes as C

import numpy as np

MAXTP1, MAXTP2, MAXTP3, MAXTP4, MAXWEI = 30, 24, 50, 20, 10
MCS_INVALID_COORD = 728492.0
MAX_LNA, MAX_NTMP, MAX_REFCONTACT = 256, 64, 256
MOVE_UNIFORM, MOVE_GAUSSIAN = 0, 1

STATUS = {0: "GD_OK", 1: "GD_ERR_INVALID", 2: "GD_ERR_CAPACITY", 3: "GD_ERR_HIP",
          4: "GD_ERR_STATE", 5: "GD_ERR_NOMEM"}

_T = C.c_float * MAXTP1 * MAXTP2


class EnePara(C.Structure):
    _fields_ = [
        ("p1a", _T), ("p2a", _T), ("lj0", C.c_float), ("lj1", C.c_float),
        ("el0", C.c_float), ("el1", C.c_float), ("ele", C.c_float * MAXTP3),
        ("a1", C.c_float), ("b1", C.c_float),
        ("pmf0", _T), ("pmf1", _T), ("hdb0", _T), ("hdb1", _T),
        ("hpp", C.c_float * MAXTP4), ("hpl0", C.c_float * MAXTP2),
        ("hpl1", C.c_float * MAXTP2), ("hpl2", C.c_float * MAXTP2),
        ("kde2", C.c_float), ("kde3", C.c_float),
        ("w", C.c_float * MAXWEI), ("a_para", C.c_float * MAXWEI),
        ("b_para", C.c_float * MAXWEI),
    ]


class Temp(C.Structure):
    _fields_ = [("t", C.c_float), ("minus_beta", C.c_float), ("order", C.c_int)]


class Replica(C.Structure):
    _fields_ = [("idx_rep", C.c_int), ("idx_prt", C.c_int), ("idx_tmp", C.c_int),
                ("idx_lig", C.c_int)]


class Energy(C.Structure):
    _fields_ = [("e", C.c_float * MAXWEI), ("cms", C.c_float), ("rmsd", C.c_float)]


class Record(C.Structure):
    _fields_ = [("replica", Replica), ("energy", Energy), ("movematrix", C.c_float * 6),
                ("step", C.c_int)]


ENERGY_DTYPE = np.dtype([("e", np.float32, (MAXWEI,)), ("cms", np.float32),
                         ("rmsd", np.float32)])
RECORD_DTYPE = np.dtype([("replica", np.int32, (4,)), ("energy", ENERGY_DTYPE),
                         ("movematrix", np.float32, (6,)), ("step", np.int32)])
assert ENERGY_DTYPE.itemsize == C.sizeof(Energy)
assert RECORD_DTYPE.itemsize == C.sizeof(Record)

_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int)


class ComplexDesc(C.Structure):
    _fields_ = [
        ("n_lig", C.c_int), ("n_prt", C.c_int), ("n_tmp", C.c_int),
        ("lna", C.c_int), ("pnp", C.c_int), ("pnk", C.c_int), ("pos", C.c_int),
        ("n_res", C.c_int), ("n_mcs", C.c_int),
        ("lig_x", _fp), ("lig_y", _fp), ("lig_z", _fp),
        ("lig_t", _ip), ("lig_c", _fp), ("lig_n", _ip),
        ("pocket_center", C.c_float * 3),
        ("prt_x", _fp), ("prt_y", _fp), ("prt_z", _fp),
        ("prt_t", _ip), ("prt_c", _ip), ("prt_ele", _fp), ("prt_seq3r", _ip),
        ("prt_c0_and_d12_or_c2", _ip), ("prt_hpp", _fp),
        ("psp", _fp),
        ("kde_x", _fp), ("kde_y", _fp), ("kde_z", _fp), ("kde_t", _ip),
        ("mcs_x", _fp), ("mcs_y", _fp), ("mcs_z", _fp), ("mcs_tcc", _fp),
        ("enepara", C.POINTER(EnePara)),
        ("temp", C.POINTER(Temp)),
        ("move_scale", C.c_float * 6),
        ("move_kind", C.c_int),
    ]


class Opts(C.Structure):
    _fields_ = [("device", C.c_int), ("seed", C.c_uint64), ("steps_per_dump", C.c_int),
                ("record", C.c_int), ("lanes_per_replica", C.c_int),
                ("atoms_per_lane", C.c_int), ("verbose", C.c_int),
                ("receptor_split", C.c_int), ("ladders_per_pack", C.c_int),
                ("launch", C.c_int), ("receptor_tile", C.c_int)]


NOALLOC_SCRATCH_BYTES = 4 << 20  # GD_NOALLOC_SCRATCH_BYTES
NOALLOC_MAX_POSES = 32768        # GD_NOALLOC_MAX_POSES

LAUNCH_NO_GRAPH = 1
LAUNCH_PERSISTENT = 2
LAUNCH_SPREAD = 4
LAUNCH_NO_SPREAD = 8


class RunStats(C.Structure):
    _fields_ = [("t_compute", C.c_double), ("t_wall", C.c_double), ("steps", C.c_int64),
                ("replica_steps", C.c_int64), ("accepted", C.c_int64),
                ("exchanges_accepted", C.c_int64), ("kernel_launches", C.c_int),
                ("kernel_ms_avg", C.c_double)]


RECORD_SINK = C.CFUNCTYPE(None, C.c_void_p, C.c_int, C.POINTER(Record), C.c_int64)


def fptr(a):
    return a.ctypes.data_as(_fp)


def iptr(a):
    return a.ctypes.data_as(_ip)


def make_desc(cx):
    """Build a ComplexDesc pointing into the numpy arrays of a synth.Complex.

    The returned object keeps references to every array it points into.
    """
    d = ComplexDesc()
    d.n_lig, d.n_prt, d.n_tmp = cx.n_lig, cx.n_prt, cx.n_tmp
    d.lna, d.pnp, d.pnk, d.pos = cx.lna, cx.pnp, cx.pnk, cx.pos
    d.n_res, d.n_mcs = cx.n_res, cx.n_mcs
    d.lig_x, d.lig_y, d.lig_z = fptr(cx.lig_x), fptr(cx.lig_y), fptr(cx.lig_z)
    d.lig_t, d.lig_c, d.lig_n = iptr(cx.lig_t), fptr(cx.lig_c), iptr(cx.lig_n)
    for i in range(3):
        d.pocket_center[i] = float(cx.pocket_center[i])
    d.prt_x, d.prt_y, d.prt_z = fptr(cx.prt_x), fptr(cx.prt_y), fptr(cx.prt_z)
    d.prt_t, d.prt_c, d.prt_ele = iptr(cx.prt_t), iptr(cx.prt_c), fptr(cx.prt_ele)
    d.prt_seq3r = iptr(cx.prt_seq3r)
    d.prt_c0_and_d12_or_c2 = iptr(cx.prt_c0_and_d12_or_c2)
    d.prt_hpp = fptr(cx.prt_hpp)
    d.psp = fptr(cx.psp)
    d.kde_x, d.kde_y, d.kde_z, d.kde_t = (fptr(cx.kde_x), fptr(cx.kde_y), fptr(cx.kde_z),
                                          iptr(cx.kde_t))
    d.mcs_x, d.mcs_y, d.mcs_z = fptr(cx.mcs_x), fptr(cx.mcs_y), fptr(cx.mcs_z)
    d.mcs_tcc = fptr(cx.mcs_tcc)
    ep = EnePara()
    C.memmove(C.addressof(ep), cx.enepara.ctypes.data, C.sizeof(EnePara))
    temps = (Temp * cx.n_tmp)()
    for i in range(cx.n_tmp):
        temps[i].t = float(cx.temp_t[i])
        temps[i].minus_beta = float(cx.temp_minus_beta[i])
        temps[i].order = i
    d.enepara = C.pointer(ep)
    d.temp = C.cast(temps, C.POINTER(Temp))
    for i in range(6):
        d.move_scale[i] = float(cx.move_scale[i])
    d.move_kind = int(cx.move_kind)
    d._keep = (cx, ep, temps)
    return d
