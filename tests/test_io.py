"""Input loaders (SDF, PDB, .ff, parameters) against the reference's own
load tests and the invariants of its host derivations. CPU only: the loader
is host code in the library.

Fixtures under tests/golden/inputs are the reference's data files
(data/1a07C1, data/edud, data/10gs, data/parameters), copied by
tests/golden/copy_inputs.sh.
"""
import math
import os

import numpy as np
import pytest

from geaxdock_gpu_amd import GeauxDockError, io

G = os.path.join(os.path.dirname(__file__), "golden", "inputs")
PARAS = os.path.join(G, "parameters", "paras")
A07 = [os.path.join(G, "1a07C1", f) for f in ("1a07C1.sdf", "1a07C.pdb", "1a07C1-0.8.ff")]
GS10 = [os.path.join(G, "10gs", f) for f in ("10gs_ligand.sdf", "10gsA00.pdb", "10gsA00.ff")]

# src/data.C:532-583 and :391-459 (type name -> code = position)
LIG_TYPES = ["Br", "C.1", "C.2", "C.3", "C.ar", "C.cat", "Cl", "F", "I", "N.1", "N.2", "N.3",
             "N.4", "N.am", "N.ar", "N.pl3", "O.2", "O.3", "O.co2", "P.3", "S.2", "S.3", "S.O",
             "S.O2"]
PNT_TYPES = ["CA", "PP", "A1", "C1", "D1", "E1", "E2", "F1", "F2", "H1", "H2", "I1", "K1", "K2",
             "L1", "M1", "M2", "N1", "P1", "Q1", "Q2", "R1", "R2", "S1", "T1", "V1", "W1", "W2",
             "Y1", "Y2"]


def almost_equal(a, b, ulps=4):
    """googletest's FloatingPoint::AlmostEquals: within 4 float32 ULPs."""
    ia = int(np.float32(a).view(np.int32))
    ib = int(np.float32(b).view(np.int32))
    ia = ia if ia >= 0 else -(ia & 0x7FFFFFFF)
    ib = ib if ib >= 0 else -(ib & 0x7FFFFFFF)
    return abs(ia - ib) <= ulps


@pytest.fixture(scope="module")
def a07():
    return io.load_complex(*A07, PARAS, complex_id="1a07C1")


def test_weight_file_matches_paras_wei():
    """load_test.C:15-30: weights read from 08ff_opt equal the WEI line of paras."""
    base = io.load_complex(*A07, PARAS, complex_id="1a07C1")
    opt = io.load_complex(*A07, PARAS, complex_id="1a07C1",
                          weight=os.path.join(G, "parameters", "08ff_opt"))
    w0 = np.array(base.enepara_struct.w[:], np.float32)
    w1 = np.array(opt.enepara_struct.w[:], np.float32)
    assert np.all(np.abs(w0 - w1) < 1e-4)
    assert np.any(w0 != 0)


def test_normalization_files_replace_noa_nob():
    cx = io.load_complex(*A07, PARAS, complex_id="1a07C1",
                         nor_a=os.path.join(G, "parameters", "08_nor_a"),
                         nor_b=os.path.join(G, "parameters", "08_nor_b"))
    with open(os.path.join(G, "parameters", "08_nor_a")) as f:
        a = [float(l[:30]) for l in f if l.strip()]
    with open(os.path.join(G, "parameters", "08_nor_b")) as f:
        b = [float(l[:30]) for l in f if l.strip()]
    n = min(len(a), 10)
    np.testing.assert_allclose(cx.enepara_struct.a_para[:n], a[:n], rtol=1e-6)
    np.testing.assert_allclose(cx.enepara_struct.b_para[:len(b[:10])], b[:10], rtol=1e-6)


def test_sdf_1a07C1():
    """load_test.C:70-110."""
    r = io.sdf_read(A07[0])
    assert r.records == 1
    assert r.ensemble_lines == r.ensemble_total == len(r.rmsd) == 20
    assert r.conformers == 21
    assert almost_equal(45.6740, r.xyz[0, 0, 0])
    assert almost_equal(-5.2246, r.xyz[2, 1, 0])  # docked conformer 2 (every RMSD > 0.1)
    assert r.typed == 1 and np.all(r.types >= 0)


def test_sdf_1b9v_4():
    """load_test.C:113-176 (the EDUD file has no OB_ATOM_TYPES)."""
    path = os.path.join(G, "edud", "1b9v_4.sdf")
    r = io.sdf_read(path)
    assert r.records == 6
    assert r.ensemble_total == 50
    assert r.ensemble_lines == 50 and len(r.rmsd) == 50
    assert r.lna == 23 and r.lnb == 23
    assert r.typed == 0 and np.all(r.types == -1)
    assert almost_equal(-0.0184, r.xyz[0, 0, 1])   # atom block x[1]
    assert almost_equal(2.1011, r.xyz[1, 0, 1])    # ENSEMBLE_COORDS line 0, x[1]
    assert almost_equal(-2.2202, r.xyz[3, 2, 1])   # ENSEMBLE_COORDS line 2, z[1]
    c = io.sdf_read(path, centre=True)
    assert almost_equal(-3.6306739, c.xyz[0, 0, 0])
    # line 0 has RMSD 0.000: it is the atom block, and the docked set skips it
    assert r.rmsd[0] == 0.0 and r.conformers == 50
    np.testing.assert_allclose(c.xyz[1], c.xyz[0], atol=2e-4)


def test_pocket_centre_1b9v_4():
    """load_test.C:168-174: loadPocketCenter of 1b9v_4.ff (loaded beside the
    1a07C1 ligand and receptor: the .ff parser does not look at them)."""
    cx = io.load_complex(A07[0], A07[1], os.path.join(G, "edud", "1b9v_4.ff"), PARAS,
                         complex_id="1b9v_4")
    assert almost_equal(-9.590, cx.pocket_center[1])


def test_sdf_1e66_4():
    """load_test.C:179-214."""
    r = io.sdf_read(os.path.join(G, "edud", "1e66_4.sdf"))
    assert r.records == 6
    assert r.ensemble_total == 50 and r.ensemble_lines == 50 and len(r.rmsd) == 50
    assert r.lna == 28 and r.lnb == 31


def test_later_records_and_bad_record():
    path = os.path.join(G, "edud", "1e66_4.sdf")
    r5 = io.sdf_read(path, record=5)
    assert r5.records == 6 and r5.lna > 0
    with pytest.raises(GeauxDockError):
        io.sdf_read(path, record=6)


def test_10gs_loads_with_one_receptor_model():
    """load_test.C:217-240: the 10gs ligand loads; 10gsA00.pdb is one model."""
    cx = io.load_complex(*GS10, PARAS, complex_id="10gsA00")
    assert cx.n_prt == 1
    assert cx.n_lig == io.sdf_read(GS10[0]).conformers


def test_1a07C_has_three_receptor_models(a07):
    """h5_test.C:40: prt_file.conf_total == 3 for 1a07C.pdb."""
    assert a07.n_prt == 3
    assert a07.n_lig == 21 and a07.lna == 34


def test_untyped_ligand_is_rejected():
    with pytest.raises(GeauxDockError, match="OB_ATOM_TYPES"):
        io.load_complex(os.path.join(G, "edud", "1b9v_4.sdf"), A07[1], A07[2], PARAS)


def test_missing_file_is_an_error():
    with pytest.raises(GeauxDockError, match="cannot open"):
        io.load_complex(A07[0] + ".none", A07[1], A07[2], PARAS)


def _heavy_centroid(sdf):
    with open(sdf) as f:
        lines = f.read().split("\n")
    lna = int(lines[3][:3])
    xyz = []
    for l in lines[4:4 + lna]:
        if l[31:34].strip() != "H":
            xyz.append([float(l[0:10]), float(l[10:20]), float(l[20:30])])
    return np.mean(np.array(xyz), axis=0)


@pytest.mark.parametrize("files", [A07, GS10], ids=["1a07C1", "10gsA00"])
def test_pocket_centre_is_heavy_atom_centroid(files):
    """The .ff CENTER line is the heavy-atom centroid of the ligand's atom
    block (data/sdf_center.py writes it that way)."""
    cx = io.load_complex(*files, PARAS)
    np.testing.assert_allclose(cx.pocket_center, _heavy_centroid(files[0]), atol=1e-4)


def test_conformers_are_centred_in_docked_order(a07):
    """Each docked conformer sits at its own centroid, and they are the atom
    block and then the ENSEMBLE_COORDS lines with RMSD > 0.1, in file order."""
    r = io.sdf_read(A07[0], centre=True)
    keep = [0] + [1 + k for k in range(r.ensemble_lines) if r.rmsd[k] > 0.1]
    lx = a07.lig_x.reshape(a07.n_lig, a07.lna)
    ly = a07.lig_y.reshape(a07.n_lig, a07.lna)
    lz = a07.lig_z.reshape(a07.n_lig, a07.lna)
    got = np.stack([lx, ly, lz], axis=1)
    np.testing.assert_array_equal(got, r.xyz[keep])
    assert np.abs(got.mean(axis=2)).max() < 1e-5
    raw = io.sdf_read(A07[0])
    np.testing.assert_allclose(got + raw.xyz[keep].mean(axis=2, keepdims=True), raw.xyz[keep],
                               atol=2e-5)
    np.testing.assert_array_equal(a07.lig_t, raw.types)
    assert a07.lig_c.dtype == np.float32 and np.any(a07.lig_c != 0)


def _paras_tables():
    vdw, pmf, hdb, hpl = {}, {}, {}, {}
    kde = None
    el = lj = None
    with open(PARAS) as f:
        for l in f:
            t = l.split()
            if len(l) <= 3 or not t:
                continue
            if t[0] in ("VDW", "PMF", "HDB"):
                d = {"VDW": vdw, "PMF": pmf, "HDB": hdb}[t[0]]
                d[(PNT_TYPES.index(t[1]), LIG_TYPES.index(t[2]))] = (float(t[3]), float(t[4]))
            elif t[0] == "HPL":
                hpl[LIG_TYPES.index(t[1])] = (float(t[2]), float(t[3]))
            elif t[0] == "KDE":
                kde = float(t[1])
            elif t[0] == "PEL":
                el = (float(t[1]), float(t[2]))
            elif t[0] == "PLJ":
                lj = (float(t[1]), float(t[2]), float(t[3]))
    return vdw, pmf, hdb, hpl, kde, el, lj


def test_enepara_derivation_invariants(a07):
    """OptimizeEnepara (src/util.C:717-768) checked by inverting it:
    sigma and epsilon come back out of the two Lennard-Jones factors, the PMF
    and H-bond tables land at [ligand][point], HPP spans [-1, 1], the HPL
    Gaussians are normalized and the electrostatic a1/b1 meet 2a1 + 3b1 = -1."""
    ep = a07.enepara_struct
    vdw, pmf, hdb, hpl, kde, el, lj = _paras_tables()
    assert len(vdw) == 720 and len(pmf) == 720
    for (pt, lt), (sig, eps) in vdw.items():
        p1, p2 = ep.p1a[lt][pt], ep.p2a[lt][pt]
        if eps == 0.0:
            assert p1 == 0.0 and p2 == 0.0
            continue
        s3 = 1.5 * p1 / p2  # (sig * lj2)^3
        s = s3 ** (1.0 / 3.0)
        assert abs(s / lj[2] - sig) <= 1e-5 * sig, (pt, lt)
        assert abs(p2 / (3.0 * s ** 6) - eps) <= 1e-4 * abs(eps), (pt, lt)
    for (pt, lt), (a, b) in pmf.items():
        assert ep.pmf0[lt][pt] == np.float32(a) and ep.pmf1[lt][pt] == np.float32(b)
    for (pt, lt), (a, b) in hdb.items():
        assert ep.hdb0[lt][pt] == np.float32(a)
        assert abs(ep.hdb1[lt][pt] * b - 1.0) < 1e-6
    hpp = np.array(ep.hpp[:])
    assert hpp.min() == -1.0 and hpp.max() == 1.0
    for lt, (h0, h1) in hpl.items():
        assert ep.hpl0[lt] == np.float32(h0) and ep.hpl1[lt] == np.float32(h1)
        assert abs(math.exp(ep.hpl2[lt]) * h1 * math.sqrt(2 * math.pi) - 1.0) < 1e-5
    assert abs(ep.kde2 * kde * kde + 0.5) < 1e-6
    assert abs(ep.kde3 - (kde * math.sqrt(2 * math.pi)) ** 3) < 1e-5 * ep.kde3
    assert ep.el0 == np.float32(el[0]) and ep.el1 == np.float32(el[1])
    assert abs(2 * ep.a1 + 3 * ep.b1 + 1.0) < 1e-6
    assert ep.lj0 == np.float32(lj[0]) and ep.lj1 == np.float32(lj[1])


def test_temperature_ladder_matches_host_derivation():
    """SetTemperature from the loader equals the engine's host ladder."""
    from geaxdock_gpu_amd import synth
    cx = io.load_complex(*A07, PARAS, n_tmp=8, floor_temp=0.044, ceiling_temp=0.036)
    t, mb = synth.temperature_ladder(8, 0.044, 0.036)
    np.testing.assert_array_equal(cx.temp_t, t)
    np.testing.assert_array_equal(cx.temp_minus_beta, mb)
    np.testing.assert_array_equal(cx.move_scale, np.float32([0.02] * 3 + [0.08] * 3))


def _pdb_blocks(path):
    """ATOM records per block ending at ENDMDL or TER, empty blocks dropped."""
    blocks, cur = [], []
    with open(path) as f:
        for l in f:
            l = l.rstrip("\n")
            if len(l) > 53 and l.startswith("ATOM  "):
                cur.append((int(l[22:26]), l[12:16], float(l[30:38]), float(l[38:46]),
                            float(l[46:54])))
            elif len(l) <= 53 and (l.startswith("ENDMDL") or l.startswith("TER")):
                if cur:
                    blocks.append(cur)
                cur = []
    return blocks


@pytest.mark.parametrize("files", [A07, GS10], ids=["1a07C1", "10gsA00"])
def test_ca_points_are_ca_atoms_per_model(files):
    """Every CA point of receptor conformer m is that residue's CA atom in
    the m-th block of the PDB (the coarse model keeps CA atoms as they are),
    and a one-centroid side-chain point is the mean of the residue's atoms
    other than N, CA, C, O."""
    cx = io.load_complex(*files, PARAS)
    blocks = _pdb_blocks(files[1])
    assert cx.n_prt == len(blocks)
    px = cx.prt_x.reshape(cx.n_prt, cx.pnp)
    py = cx.prt_y.reshape(cx.n_prt, cx.pnp)
    pz = cx.prt_z.reshape(cx.n_prt, cx.pnp)
    ca_pts = np.flatnonzero(cx.prt_t == 0)
    assert len(ca_pts) == cx.residues
    for m, b in enumerate(blocks):
        ca = np.array([[x, y, z] for _, nm, x, y, z in b if nm == " CA "], np.float32)
        assert len(ca) == len(ca_pts)
        got = np.stack([px[m, ca_pts], py[m, ca_pts], pz[m, ca_pts]], axis=1)
        np.testing.assert_array_equal(got, ca)
    # ALA's one side-chain point (A1, code 2) is its CB
    a1 = np.flatnonzero(cx.prt_t == 2)
    if len(a1):
        seqs = [s for s, nm, *_ in blocks[0] if nm == " CA "]
        i = a1[0]
        r = int(np.searchsorted(ca_pts, i, side="right")) - 1  # residue of the point
        side = [(x, y, z) for s, nm, x, y, z in blocks[0]
                if s == seqs[r] and nm not in (" N  ", " CA ", " C  ", " O  ")]
        np.testing.assert_allclose([px[0, i], py[0, i], pz[0, i]], np.mean(side, axis=0),
                                   atol=1e-4)


@pytest.mark.parametrize("line,msg", [("PSP 1000000000 C.3 0.5", "PSP residue"),
                                      ("MCS 1a07C1 0.5 1 99999 1.0 2.0 3.0", "MCS field")])
def test_out_of_range_indices_in_ff_are_errors(tmp_path, line, msg):
    """Indices that size host tables are bounded: a corrupt .ff line is an
    error, not a huge allocation."""
    ff = tmp_path / "bad.ff"
    ff.write_text(open(A07[2]).read().rstrip("\n") + "\n" + line + "\n")
    with pytest.raises(GeauxDockError, match=msg):
        io.load_complex(A07[0], A07[1], str(ff), PARAS, complex_id="1a07C1")


def test_bad_temperature_ladder_is_an_error():
    with pytest.raises(GeauxDockError, match="temperature"):
        io.load_complex(*A07, PARAS, n_tmp=4, floor_temp=0.3, ceiling_temp=0.0)
