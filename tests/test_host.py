"""CPU tests of the C-ABI libraries and the host-side setup code: every
declared entry point is exported, and the host's potential reader,
Chebyshev fit, lattice, seeded inputs, gz files and grid agree with the
compiled reference's fixtures and the oracle.  No GPU is needed."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import comd_amd as CA
import oracle_py as O
from helpers import GOLDEN, POTS, ROOT


def declared_functions(header):
    """Names of the functions an include/*.h file declares."""
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    names = set()
    for m in re.finditer(r"^[A-Za-z_][\w \t\*]*?\b([A-Za-z_]\w*)\s*\(([^;{]*?)\)\s*;", text, flags=re.M):
        if m.group(1) not in ("if", "while", "for", "return", "sizeof"):
            names.add(m.group(1))
    return sorted(names)


@pytest.mark.parametrize("header,lib", [("comd_hip.h", 0), ("comd_host.h", 1), ("comd_dropin.h", 2)])
def test_libraries_export_every_declared_function(header, lib):
    names = declared_functions(header)
    assert len(names) > (10 if lib < 2 else 4)
    handle = (CA.libraries() + (CA.dropin(),))[lib]
    missing = [n for n in names if not hasattr(handle, n)]
    assert not missing, f"{header}: not exported by the library: {missing}"


def test_c_abi_has_no_torch_types():
    for h in ("comd_hip.h", "comd_host.h"):
        text = open(os.path.join(ROOT, "include", h)).read()
        for token in ("at::Tensor", "c10::", "torch::", "#include <torch"):
            assert token not in text


def test_host_tables_and_chebyshev_match_reference(cu500):
    g = np.load(os.path.join(GOLDEN, "cu500_pot.npz"))
    assert (cu500.cutoff, cu500.mass, cu500.lat) == (float(g["cutoff"]), float(g["mass"]), float(g["lat"]))
    for t, name in zip(cu500.tables, ("phi", "rho", "F")):
        n, x0, xn, invdx = g[f"{name}_hdr"]
        assert (t["n"], t["x0"], t["xn"], t["invdx"]) == (int(n), x0, xn, invdx)
        np.testing.assert_array_equal(t["values"][:4], g[f"{name}_head"])
        np.testing.assert_array_equal(t["values"][-4:], g[f"{name}_tail"])
    ref = g["cheb"]
    scale = np.abs(ref).max(axis=1, keepdims=True)
    assert (np.abs(cu500.cheb - ref) <= 1e-14 * scale).all()
    np.testing.assert_array_equal(cu500.cheb_dom, g["cheb_ab"])


def test_host_tables_equal_oracle_tables(cu500, ocu500):
    for k in range(3):
        o = ocu500.table(k)
        np.testing.assert_array_equal(cu500.tables[k]["values"], o["values"])


def test_fcc_positions_follow_create_fcc_lattice():
    """Binning the host lattice gives exactly the oracle's create_fcc_lattice
    state (same ids, cells, slots and relative positions)."""
    pot = O.OraclePot.eam(POTS, "cu500")
    for n, lat, dg in ((5, 3.615, 1.0), (6, 3.615, 1.07)):
        xyz = CA.fcc_positions(n, n, n, lat, dg)
        assert xyz.shape == (4 * n ** 3, 3)
        a = O.OracleSim.fcc(pot, n, lat, defgrad=dg)
        bounds = a.info()["bounds"]
        b = O.OracleSim.from_atoms(pot, tuple(bounds), xyz)
        sa, sb = a.state(), b.state()
        for k in ("natoms", "id", "r"):
            np.testing.assert_array_equal(sa[k], sb[k])


def test_seeded_state_statistics_and_determinism():
    n = 8
    xyz0 = CA.fcc_positions(n, n, n, 3.615)
    x1, p1 = CA.seed_state(xyz0, 63.55, 300.0, 0.05, 11)
    x2, p2 = CA.seed_state(xyz0, 63.55, 300.0, 0.05, 11)
    x3, _ = CA.seed_state(xyz0, 63.55, 300.0, 0.05, 12)
    np.testing.assert_array_equal(x1, x2)
    np.testing.assert_array_equal(p1, p2)
    assert not np.array_equal(x1, x3)
    rmass = CA.AMU_TO_ME * 63.55
    assert np.abs(p1.sum(0)).max() < 1e-9 * np.abs(p1).sum()
    temp = (p1 ** 2).sum() / rmass / (3 * len(p1) * CA.KB_HARTREE)
    assert abs(temp / 300.0 - 1) < 0.06
    disp = np.sqrt(((x1 - xyz0) ** 2).mean())
    assert abs(disp / 0.05 - 1) < 0.06


def test_gz_round_trip_is_exact(tmp_path):
    n = 4
    xyz, p = CA.seed_state(CA.fcc_positions(n, n, n, 3.615), 63.55, 300.0, 0.05, 3)
    bounds = (n * 3.615,) * 3
    path = tmp_path / "s.inp.gz"
    CA.write_gz(path, bounds, xyz, p)
    b, x2, p2 = CA.read_gz(path)
    assert tuple(b) == bounds
    np.testing.assert_array_equal(x2, xyz)
    np.testing.assert_array_equal(p2, p)
    # and the oracle's reader (src-flat/read.c:155-241) sees the same atoms
    pot = O.OraclePot.eam(POTS, "cu500")
    a = O.OracleSim.gz(pot, path)
    bsim = O.OracleSim.from_atoms(pot, bounds, xyz, p)
    sa, sb = a.state(), bsim.state()
    for k in ("natoms", "id", "r", "p"):
        np.testing.assert_array_equal(sa[k], sb[k])


def test_read_reference_input_files():
    b, xyz, p = CA.read_gz(os.path.join(GOLDEN, "data", "2.inp.gz"))
    assert tuple(b) == (40.0, 40.0, 40.0) and xyz.shape == (2, 3)
    np.testing.assert_array_equal(xyz[0], [1.0, 39.5, 0.1])
    b, xyz, p = CA.read_gz(os.path.join(GOLDEN, "data", "8k.inp.gz"))
    assert xyz.shape[0] == 8192  # first line of the file: "8192 ..."


@pytest.mark.parametrize("bounds,defgrad", [((21.69, 21.69, 21.69), 1.0), ((23.2083, 21.69, 21.69), 1.0),
                                            ((144.6, 144.6, 144.6), 1.0), ((40.0, 40.0, 40.0), 0.0),
                                            ((21.69, 21.69, 21.69), 1.07)])
def test_grid_matches_alloc_domains(cu500, bounds, defgrad):
    """comd_grid_for restates allocDomains (src-flat/domains.c:155-200)."""
    hip, _ = CA.libraries()
    prm = CA.Params()
    cu500.fill(prm)
    assert hip.comd_grid_for(C.byref(prm), (C.c_double * 3)(*bounds), defgrad, 1.0) == 0
    pot = O.OraclePot.eam(POTS, "cu500")
    sim = O.OracleSim(O.lib().orc_sim_create(pot.h, (C.c_double * 3)(*bounds), defgrad, 1.0, 1), pot)
    info = sim.info()
    assert tuple(prm.nbx) == tuple(info["nbx"])
    np.testing.assert_array_equal(np.array(prm.boxsize[:]), info["boxsize"])


def test_grid_rejects_box_smaller_than_cutoff(cu500):
    hip, _ = CA.libraries()
    prm = CA.Params()
    cu500.fill(prm)
    assert hip.comd_grid_for(C.byref(prm), (C.c_double * 3)(3.0, 20.0, 20.0), 1.0, 1.0) == 6


def test_struct_layouts_match_reference():
    """The structs that cross the drop-in boundary have the reference's
    size and field offsets (fixture: oracle/ref_layout.c compiled against
    the reference's headers; here against include/comd_host.h)."""
    import json
    import subprocess
    import tempfile

    ref = json.load(open(os.path.join(GOLDEN, "ref_layout.json")))
    with tempfile.TemporaryDirectory() as tmp:
        exe = os.path.join(tmp, "layout")
        subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), "-o", exe,
                        os.path.join(ROOT, "oracle", "ref_layout.c")], check=True)
        ours = json.loads(subprocess.run([exe], check=True, capture_output=True, text=True).stdout)
    assert ours == ref


def test_parse_command_line_keeps_reference_command_t():
    """parseCommandLine fills the reference's command_t (mycommand.h:60-75)
    and returns nothing, as the reference's does; this library's extra
    options go to the driver settings, not into command_t."""
    _, host = CA.libraries()
    host.comd_host_run_options.restype = C.c_void_p
    args = [b"comd", b"-e", b"-p", b"cu500", b"-x", b"5", b"-y", b"6", b"-z", b"7", b"-t", b"300", b"-s", b"1.05",
            b"--precision", b"mixed", b"--nsteps", b"4", b"--seed", b"9"]
    argv = (C.c_char_p * len(args))(*args)
    cmd = CA.CommandT()
    assert C.sizeof(cmd) == json_size("command_t")
    assert host.parseCommandLine(C.byref(cmd), len(args), argv) is None
    assert host.comd_host_status() == 0
    assert (cmd.doeam, cmd.nx, cmd.ny, cmd.nz, cmd.periodic) == (1, 5, 6, 7, 1)
    assert (cmd.temp, cmd.defgrad, cmd.bf) == (300.0, 1.05, 1.0)
    assert cmd.potname == b"cu500" and cmd.potdir == b"pots"
    # driver settings: precision, cheby, seed, disp, niter, nsteps, ...
    opts = (C.c_int * 2).from_address(host.comd_host_run_options())
    assert opts[0] == CA.PREC_MIXED
    args = [b"comd", b"--bogus"]
    host.parseCommandLine(C.byref(cmd), 2, (C.c_char_p * 2)(*args))
    assert host.comd_host_status() != 0


def json_size(name):
    import json

    return json.load(open(os.path.join(GOLDEN, "ref_layout.json")))[name]["size"]


def test_create_fcc_lattice_reference_semantics():
    """create_fcc_lattice takes a reference-layout command_t by value and
    builds the reference's 0 K lattice (ic_fcc.c:63-141: type 1, ids in
    creation order, atoms at rest; cmd.temp unused), the oracle's state."""
    _, host = CA.libraries()
    pot = host.setEamPot(POTS.encode(), b"cu500")
    cmd = CA.CommandT()
    cmd.nx, cmd.ny, cmd.nz = 5, 5, 5
    cmd.lat, cmd.defgrad, cmd.bf, cmd.temp = 3.615, 1.07, 1.0, 600.0
    cmd.periodic = 1
    sp = host.create_fcc_lattice(cmd, C.cast(pot, C.c_void_p))
    assert sp
    s = sp.contents
    nb = s.nboxes
    natoms = np.ctypeslib.as_array(s.natoms, shape=(nb,)).copy()
    ids = np.ctypeslib.as_array(s.id, shape=(nb * 64,)).copy()
    typ = np.ctypeslib.as_array(s.iType, shape=(nb * 64,)).copy()
    r = np.ctypeslib.as_array(s.r, shape=(nb * 64, 3)).copy()
    p = np.ctypeslib.as_array(s.p, shape=(nb * 64, 3)).copy()
    o = O.OracleSim.fcc(O.OraclePot.eam(POTS, "cu500"), 5, 3.615, defgrad=1.07)
    so = o.state()
    m = (np.arange(64)[None, :] < natoms[:, None]).reshape(-1)
    np.testing.assert_array_equal(natoms, so["natoms"])
    np.testing.assert_array_equal(ids[m], so["id"][m])
    np.testing.assert_array_equal(r[m], so["r"][m])
    assert (typ[m] == 1).all()
    assert (p[m] == 0.0).all()
    assert s.ntot == 500 and tuple(s.bounds) == tuple(o.info()["bounds"])
    host.destroySimulation(C.byref(sp))


def test_dropin_library_exports_and_links_no_reference_names():
    """libcomd_dropin.so exports what include/comd_dropin.h declares and
    defines none of the reference's function names (so a reference build
    can link it beside its own objects)."""
    import subprocess

    names = declared_functions("comd_dropin.h")
    assert {"comd_hip_eam_force", "comd_hip_lj_force", "comd_hip_nTimeSteps", "comd_hip_reBoxAll"} <= set(names)
    d = CA.dropin()
    assert all(hasattr(d, n) for n in names)
    syms = subprocess.run(["nm", "-D", "--defined-only", os.path.join(ROOT, "comd_amd", "libcomd_dropin.so")],
                          check=True, capture_output=True, text=True).stdout.split()
    reference_names = {"eamForce", "LJ", "computeForce", "nTimeSteps", "reBoxAll", "setEamPot", "getLJPot",
                       "allocDomains", "putAtomInBox", "create_fcc_lattice", "fromFileGzip", "parseCommandLine",
                       "initSimFromCmdLine", "do_compute_work", "setChebPot", "blankSimulation"}
    assert not reference_names & set(syms)


def _simflat_from_fixture(host, fx, m):
    """A host simflat_t on the grid of data/cu6.inp.gz (fromFileGzip) holding
    state m of a reference fixture."""
    pot = host.setEamPot(POTS.encode(), b"cu500")
    cmd = CA.CommandT()
    cmd.filename = os.path.join(GOLDEN, "data", "cu6.inp.gz").encode()
    cmd.bf, cmd.defgrad, cmd.periodic = 1.0, 1.0, 1
    sp = host.fromFileGzip(cmd, C.cast(pot, C.c_void_p))
    s = sp.contents
    nb = s.nboxes
    natoms = np.ctypeslib.as_array(s.natoms, shape=(nb,))
    ids = np.ctypeslib.as_array(s.id, shape=(nb * 64,))
    r = np.ctypeslib.as_array(s.r, shape=(nb * 64, 3))
    p = np.ctypeslib.as_array(s.p, shape=(nb * 64, 3))
    na = fx[f"s{m}_natoms"]
    natoms[:] = na
    mask = (np.arange(64)[None, :] < na[:, None]).reshape(-1)
    ids[mask] = fx[f"s{m}_id"]
    r[mask] = fx[f"s{m}_r"]
    p[mask] = fx[f"s{m}_p"]
    return sp


def test_writers_match_reference_output(tmp_path):
    """writeClsman and gzASCIIWrite (src-flat/write.c:112-220) on the state
    the reference reached after 10 steps of data/cu6.inp.gz: the clsman file
    is byte-identical to the reference's, the gz text identical to its
    text; comd_read_clsman reads the atoms back exactly."""
    import gzip

    _, host = CA.libraries()
    host.writeClsman.argtypes = [C.POINTER(CA.SimflatT), C.c_char_p]
    host.writeClsman.restype = None
    host.gzASCIIWrite.argtypes = [C.POINTER(CA.SimflatT), C.c_char_p]
    host.gzASCIIWrite.restype = None
    host.comd_read_clsman.argtypes = [C.c_char_p, C.POINTER(C.c_int), C.c_double * 3,
                                      C.POINTER(C.POINTER(C.c_double)), C.POINTER(C.POINTER(C.c_double))]
    fx = np.load(os.path.join(GOLDEN, "cu6_table.npz"))
    sp = _simflat_from_fixture(host, fx, 2)
    out_b = tmp_path / "final.clsman"
    out_t = tmp_path / "final.gz"
    host.writeClsman(sp, str(out_b).encode())
    host.gzASCIIWrite(sp, str(out_t).encode())
    assert host.comd_host_status() == 0
    assert out_b.read_bytes() == open(os.path.join(GOLDEN, "cu6_final.clsman"), "rb").read()
    assert gzip.open(out_t).read() == open(os.path.join(GOLDEN, "cu6_final.txt"), "rb").read()
    # read back: absolute positions in id order, as fromFileGzip would place them
    n = C.c_int()
    b = (C.c_double * 3)()
    xp, pp = C.POINTER(C.c_double)(), C.POINTER(C.c_double)()
    assert host.comd_read_clsman(str(out_b).encode(), C.byref(n), b, C.byref(xp), C.byref(pp)) == 0
    xyz = np.ctypeslib.as_array(xp, shape=(n.value, 3)).copy()
    pm = np.ctypeslib.as_array(pp, shape=(n.value, 3)).copy()
    s = sp.contents
    bs = np.array(s.boxsize[:])
    nbx = np.array(s.nbx[:])
    na = fx["s2_natoms"]
    cells = np.repeat(np.arange(len(na)), na)
    corner = np.stack([cells % nbx[0], (cells // nbx[0]) % nbx[1], cells // (nbx[0] * nbx[1])], 1) * bs
    order = np.argsort(fx["s2_id"])
    np.testing.assert_array_equal(xyz, (corner + fx["s2_r"])[order])
    np.testing.assert_array_equal(pm, fx["s2_p"][order])
    host.destroySimulation(C.byref(sp))
