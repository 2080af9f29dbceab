"""GeauxDock input files (SDF, PDB, .ff, parameters) as an engine complex.

`load_complex` calls the library's C++ loader (include/geauxdock_io.h), which
follows the reference's `dock` driver: loadLigand, loadProtein, loadLHM,
loadEnePara, the Optimize* derivations and SetTemperature
(src/dock.C:104-186). The result carries the same fields as a synthetic
complex (`synth.Complex`), so every engine and surrogate entry point takes it.
"""
import ctypes as C

import numpy as np

from . import abi
from .synth import Complex


class InputFiles(C.Structure):
    _fields_ = [("ligand_sdf", C.c_char_p), ("protein_pdb", C.c_char_p), ("lhm_ff", C.c_char_p),
                ("enepara", C.c_char_p), ("complex_id", C.c_char_p), ("weight", C.c_char_p),
                ("nor_a", C.c_char_p), ("nor_b", C.c_char_p)]


class DockParams(C.Structure):
    _fields_ = [("n_tmp", C.c_int), ("floor_temp", C.c_float), ("ceiling_temp", C.c_float),
                ("ts", C.c_float), ("rs", C.c_float), ("move_kind", C.c_int)]


class SdfRecord(C.Structure):
    _fields_ = [("records", C.c_int), ("lna", C.c_int), ("lnb", C.c_int), ("typed", C.c_int),
                ("ensemble_total", C.c_int), ("ensemble_lines", C.c_int),
                ("conformers", C.c_int)]


def _lib():
    from . import load_library
    lib = load_library()
    if not getattr(lib, "_io_bound", False):
        lib.gd_load_complex.argtypes = [C.POINTER(InputFiles), C.POINTER(DockParams),
                                        C.POINTER(C.c_void_p), C.c_char_p, C.c_size_t]
        lib.gd_load_complex.restype = C.c_int
        lib.gd_loaded_desc.argtypes = [C.c_void_p]
        lib.gd_loaded_desc.restype = C.POINTER(abi.ComplexDesc)
        lib.gd_loaded_info.argtypes = [C.c_void_p] + [C.POINTER(C.c_int)] * 4
        lib.gd_loaded_info.restype = None
        lib.gd_loaded_free.argtypes = [C.c_void_p]
        lib.gd_loaded_free.restype = None
        lib.gd_sdf_read.argtypes = [C.c_char_p, C.c_int, C.c_int, C.POINTER(SdfRecord),
                                    C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_void_p,
                                    C.c_void_p, C.c_char_p, C.c_size_t]
        lib.gd_sdf_read.restype = C.c_int
        E = [C.c_char_p, C.c_size_t]
        lib.gd_write_states_csv.argtypes = [C.c_char_p, C.c_void_p, C.c_int64] + E
        lib.gd_write_medoids_csv.argtypes = [C.c_char_p, C.c_void_p, C.c_void_p, C.c_int64] + E
        lib.gd_records_write.argtypes = [C.c_char_p, C.c_int, C.c_void_p, C.c_int64] + E
        lib.gd_records_count.argtypes = [C.c_char_p, C.POINTER(C.c_int64)] + E
        lib.gd_records_read.argtypes = [C.c_char_p, C.c_int64, C.c_int64, C.c_void_p] + E
        lib.gd_pose_coords.argtypes = [C.POINTER(abi.ComplexDesc), C.c_int, C.c_void_p,
                                       C.c_void_p, C.c_void_p, C.c_void_p]
        lib.gd_write_pose_sdf.argtypes = [C.c_char_p, C.c_int, C.c_void_p, C.c_void_p,
                                          C.c_void_p, C.c_char_p] + E
        for f in ("gd_write_states_csv", "gd_write_medoids_csv", "gd_records_write",
                  "gd_records_count", "gd_records_read", "gd_pose_coords", "gd_write_pose_sdf"):
            getattr(lib, f).restype = C.c_int
        lib._io_bound = True
    return lib


def _b(s):
    return None if s is None else str(s).encode()


def load_complex(sdf, pdb, ff, paras, complex_id="", n_tmp=1, floor_temp=0.3, ceiling_temp=0.3,
                 ts=0.02, rs=0.08, move_kind=abi.MOVE_UNIFORM, weight=None, nor_a=None,
                 nor_b=None):
    """Load one complex. Defaults are the reference driver's (src/dock.C:28-37):
    one temperature of 0.3, translation scale 0.02, rotation scale 0.08,
    uniform moves."""
    lib = _lib()
    f = InputFiles(_b(sdf), _b(pdb), _b(ff), _b(paras), _b(complex_id), _b(weight), _b(nor_a),
                   _b(nor_b))
    p = DockParams(int(n_tmp), float(floor_temp), float(ceiling_temp), float(ts), float(rs),
                   int(move_kind))
    h = C.c_void_p()
    err = C.create_string_buffer(512)
    st = lib.gd_load_complex(C.byref(f), C.byref(p), C.byref(h), err, 512)
    if st != 0:
        from . import GeauxDockError
        raise GeauxDockError(st, err.value.decode())
    try:
        d = lib.gd_loaded_desc(h).contents
        info = [C.c_int(0) for _ in range(4)]
        lib.gd_loaded_info(h, *[C.byref(i) for i in info])
        cx = _from_desc(d)
        cx.sdf_sections, cx.ensemble_total, cx.residues, cx.kde_dropped = (i.value for i in info)
        cx.complex_id = complex_id
    finally:
        lib.gd_loaded_free(h)
    return cx


class SdfData:
    """One SDF record: counts, raw or centred conformers, RMSDs, types, charges."""


def sdf_read(path, record=0, centre=False):
    """One record of an SDF file as the file holds it (`gd_sdf_read`).

    `xyz` is (1 + ensemble_lines, 3, lna): the atom block, then every
    ENSEMBLE_COORDS line; `centre` moves each to its own centroid."""
    lib = _lib()
    info = SdfRecord()
    err = C.create_string_buffer(512)

    def call(xyz, nconf, rmsd, nr, types, charges):
        st = lib.gd_sdf_read(_b(path), int(record), int(bool(centre)), C.byref(info), xyz, nconf,
                             rmsd, nr, types, charges, err, 512)
        if st != 0:
            from . import GeauxDockError
            raise GeauxDockError(st, err.value.decode())

    call(None, 0, None, 0, None, None)
    n = 1 + info.ensemble_lines
    xyz = np.zeros((n, 3, info.lna), np.float32)
    rmsd = np.zeros(info.ensemble_lines, np.float32)
    types = np.zeros(info.lna, np.int32)
    charges = np.zeros(info.lna, np.float32)
    call(xyz.ctypes.data, n, rmsd.ctypes.data, len(rmsd), types.ctypes.data, charges.ctypes.data)
    out = SdfData()
    for f, _ in SdfRecord._fields_:
        setattr(out, f, getattr(info, f))
    out.xyz, out.rmsd, out.types, out.charges = xyz, rmsd, types, charges
    return out


def _from_desc(d):
    """Copy a descriptor's arrays into a synth.Complex-shaped object."""
    cx = Complex()
    cx.n_lig, cx.n_prt, cx.n_tmp = d.n_lig, d.n_prt, d.n_tmp
    cx.lna, cx.pnp, cx.pnk, cx.pos = d.lna, d.pnp, d.pnk, d.pos
    cx.n_res, cx.n_mcs = d.n_res, d.n_mcs

    def fa(ptr, n):
        return np.ctypeslib.as_array(ptr, shape=(n,)).astype(np.float32).copy() if n else \
            np.zeros(0, np.float32)

    def ia(ptr, n):
        return np.ctypeslib.as_array(ptr, shape=(n,)).astype(np.int32).copy() if n else \
            np.zeros(0, np.int32)

    nl, nlp, npp = d.lna, d.n_lig * d.lna, d.n_prt * d.pnp
    cx.lig_x, cx.lig_y, cx.lig_z = fa(d.lig_x, nlp), fa(d.lig_y, nlp), fa(d.lig_z, nlp)
    cx.lig_t, cx.lig_c, cx.lig_n = ia(d.lig_t, nl), fa(d.lig_c, nl), ia(d.lig_n, nl)
    cx.pocket_center = np.array(d.pocket_center[:], np.float32)
    cx.prt_x, cx.prt_y, cx.prt_z = fa(d.prt_x, npp), fa(d.prt_y, npp), fa(d.prt_z, npp)
    cx.prt_t, cx.prt_c = ia(d.prt_t, d.pnp), ia(d.prt_c, d.pnp)
    cx.prt_ele, cx.prt_seq3r = fa(d.prt_ele, d.pnp), ia(d.prt_seq3r, d.pnp)
    cx.prt_c0_and_d12_or_c2 = ia(d.prt_c0_and_d12_or_c2, d.pnp)
    cx.prt_hpp = fa(d.prt_hpp, d.pnp)
    cx.psp = fa(d.psp, abi.MAXTP2 * d.n_res)
    cx.kde_x, cx.kde_y, cx.kde_z = fa(d.kde_x, d.pnk), fa(d.kde_y, d.pnk), fa(d.kde_z, d.pnk)
    cx.kde_t = ia(d.kde_t, d.pnk)
    nm = d.pos * d.n_mcs
    cx.mcs_x, cx.mcs_y, cx.mcs_z = fa(d.mcs_x, nm), fa(d.mcs_y, nm), fa(d.mcs_z, nm)
    cx.mcs_tcc = fa(d.mcs_tcc, d.pos)
    ep = abi.EnePara()
    C.memmove(C.addressof(ep), C.cast(d.enepara, C.c_void_p).value, C.sizeof(abi.EnePara))
    cx.enepara_struct = ep
    cx.enepara = np.frombuffer(bytes(ep), dtype=np.uint8).copy()
    temps = [d.temp[i] for i in range(d.n_tmp)]
    cx.temp_t = np.array([t.t for t in temps], np.float32)
    cx.temp_minus_beta = np.array([t.minus_beta for t in temps], np.float32)
    cx.move_scale = np.array(d.move_scale[:], np.float32)
    cx.move_kind = d.move_kind
    cx.seed = None
    return cx


def _call(fn, *args):
    err = C.create_string_buffer(512)
    st = fn(*args, err, 512)
    if st != 0:
        from . import GeauxDockError
        raise GeauxDockError(st, err.value.decode())


def _records(recs):
    return np.ascontiguousarray(recs, abi.RECORD_DTYPE)


def write_states_csv(path, records):
    """printStates of accepted steps (src/util.C:1504-1550)."""
    r = _records(records)
    _call(_lib().gd_write_states_csv, _b(path), r.ctypes.data, len(r))


def write_medoids_csv(path, records, cluster_size):
    """printStates of medoids (src/util.C:1464-1502)."""
    r = _records(records)
    cs = np.ascontiguousarray(cluster_size, np.int32)
    if len(cs) != len(r):
        raise ValueError("one cluster size per record")
    _call(_lib().gd_write_medoids_csv, _b(path), r.ctypes.data, cs.ctypes.data, len(r))


class RecordFile:
    """Flat record file (gd_records_*): a 64-byte header and gd_record
    entries in append order. `sink` fits Engine.run(sink=...)."""

    HEADER = 64

    def __init__(self, path, append=False):
        self.path = path
        if not append:
            self.write([], append=False)

    def write(self, records, append=True):
        r = _records(records)
        _call(_lib().gd_records_write, _b(self.path), int(bool(append)), r.ctypes.data, len(r))

    def sink(self, cid, records):
        self.write(records)

    def __len__(self):
        n = C.c_int64(0)
        _call(_lib().gd_records_count, _b(self.path), C.byref(n))
        return n.value

    def read(self, first=0, n=None):
        if n is None:
            n = len(self) - first
        out = np.zeros(n, abi.RECORD_DTYPE)
        _call(_lib().gd_records_read, _b(self.path), int(first), int(n), out.ctypes.data)
        return out

    def memmap(self):
        """The records as a read-only numpy view of the file."""
        return np.memmap(self.path, dtype=abi.RECORD_DTYPE, mode="r", offset=self.HEADER,
                         shape=(len(self),))


def pose_coords(cx, idx_lig, movematrix):
    """PlaceLigand (src/util.C:456-497): (lna, 3) coordinates of conformer
    idx_lig under the move vector."""
    d = cx.desc()
    mm = np.ascontiguousarray(movematrix, np.float32).reshape(6)
    x, y, z = (np.zeros(cx.lna, np.float32) for _ in range(3))
    st = _lib().gd_pose_coords(C.byref(d), int(idx_lig), mm.ctypes.data, x.ctypes.data,
                               y.ctypes.data, z.ctypes.data)
    if st != 0:
        from . import GeauxDockError
        raise GeauxDockError(st, "gd_pose_coords: bad arguments")
    return np.stack([x, y, z], axis=1)


def write_pose_sdf(ligand_sdf, coords, out_path):
    """replaceLigandCoords (src/util.C:504-549) written to out_path."""
    c = np.ascontiguousarray(coords, np.float32).reshape(-1, 3)
    x, y, z = (np.ascontiguousarray(c[:, i]) for i in range(3))
    _call(_lib().gd_write_pose_sdf, _b(ligand_sdf), len(c), x.ctypes.data, y.ctypes.data,
          z.ctypes.data, _b(out_path))


def read_trace(path):
    """read2D of a trace file: rows of whitespace-separated numbers (a header
    row that is not numeric is skipped)."""
    rows = []
    with open(path) as f:
        for line in f:
            t = line.split()
            if not t:
                continue
            try:
                rows.append([float(v) for v in t])
            except ValueError:
                continue
    return rows


def trace(cx, ligand_sdf, rows, prefix):
    """The trace tool (src/trace.C:90-119): for each row (protein conformer,
    ligand conformer, six move values, ...), the ligand conformer placed by
    the move vector and written as <prefix>_<row>.sdf. Returns the paths."""
    out = []
    for i, r in enumerate(rows):
        if len(r) < 8:
            raise ValueError(f"trace row {i} has {len(r)} values, needs 8")
        path = f"{prefix}_{i}.sdf"
        write_pose_sdf(ligand_sdf, pose_coords(cx, int(r[1]), r[2:8]), path)
        out.append(path)
    return out
