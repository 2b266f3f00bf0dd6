"""GPU parity tests: the HIP path through the C-ABI against the CPU oracle
(a bit-exact restatement of src-flat, itself pinned to the compiled
reference by tests/test_oracle.py) on the same seeded inputs.

Bars: cell assignment, slot order, binned/drifted positions: bit-exact.
FP64 per-atom forces, energies, densities, F': relative 1e-10 (force error
per atom is |df| / max(|f|, rms|f|)).  Mixed precision: stated and bounded
per quantity.
"""
import os

import numpy as np
import pytest

import comd_amd as CA
import oracle_py as O
from helpers import DT_INTERNAL, GOLDEN, POTS, compact, force_rel_error, scalar_rel_error, seeded_atoms

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-10


def make_pair(pot, opot, xyz, p, bounds, evaluation="table", precision="fp64", periodic=True, bf=1.0):
    """GPU simulation and oracle simulation from the same atom list, with the
    file-input semantics (grid from the bounds, ids in list order, type 0)."""
    sim = CA.Simulation(pot, bounds, defgrad=0.0, bf=bf, periodic=periodic, precision=precision,
                        evaluation=evaluation)
    sim.bin_atoms(xyz, p, itype=0)
    osim = O.OracleSim.from_atoms(opot, bounds, xyz, p, bf=bf, periodic=periodic, itype=0)
    assert osim.info()["err"] == 0
    assert tuple(osim.info()["nbx"]) == sim.nbx
    return sim, osim


def assert_same_layout(g, o, what=""):
    assert np.array_equal(g["natoms"], o["natoms"]), f"{what}: cell counts differ"
    gc, oc = compact(g), compact(o)
    assert np.array_equal(gc["id"], oc["id"]), f"{what}: slot order differs"
    assert np.array_equal(gc["itype"], oc["itype"]), f"{what}: types differ"
    return gc, oc


def check_forces(g, o, tol, what, lj=False):
    gc, oc = assert_same_layout(g, o, what)
    ef = force_rel_error(gc["f"], oc["f"])
    ee = scalar_rel_error(gc["f"][:, 3], oc["f"][:, 3])
    print(f"[{what}] atoms={len(ef)} max force rel err {ef.max():.3e} max energy rel err {ee.max():.3e}", flush=True)
    assert ef.max() <= tol, f"{what}: force error {ef.max():.3e}"
    assert ee.max() <= tol, f"{what}: energy error {ee.max():.3e}"
    if not lj:
        er = scalar_rel_error(gc["rho"], oc["rho"])
        ei = scalar_rel_error(gc["fi"], oc["fi"])
        assert er.max() <= tol, f"{what}: rho error {er.max():.3e}"
        assert ei.max() <= tol, f"{what}: F' error {ei.max():.3e}"
    return ef.max()


# ---------------------------------------------------------------------------
# binning and rebinning: bit-exact

@pytest.mark.parametrize("n,defgrad,seed", [(6, 1.0, 3), (8, 1.07, 11), (9, 1.0, 5)])
def test_binning_bit_exact(gpu, cu500, ocu500, n, defgrad, seed):
    print(f"[case] binning n={n} defgrad={defgrad} seed={seed}", flush=True)
    xyz, p, bounds = seeded_atoms(n, 3.615, 63.55, temp=600.0, disp=0.2, seed=seed, defgrad=defgrad)
    sim, osim = make_pair(cu500, ocu500, xyz, p, bounds)
    g, o = sim.download(), osim.state()
    gc, oc = assert_same_layout(g, o, "bin")
    assert np.array_equal(gc["r"], oc["r"])
    assert np.array_equal(gc["p"], oc["p"])
    sim.rebox()
    osim.rebox()
    gc, oc = assert_same_layout(sim.download(), osim.state(), "rebox")
    assert np.array_equal(gc["r"], oc["r"])


def test_drift_rebox_bit_exact(gpu, cu500, ocu500):
    """advancePositions + reBoxAll on identical inputs, with large drifts so
    that many atoms change cells in every pass."""
    xyz, p, bounds = seeded_atoms(8, 3.615, 63.55, temp=3000.0, disp=0.1, seed=21)
    sim, osim = make_pair(cu500, ocu500, xyz, p, bounds)
    moved = 0
    for k in range(6):
        print(f"[case] drift pass {k}", flush=True)
        before = compact(osim.state())["id"].copy()
        sim.advance_positions(40 * DT_INTERNAL)
        osim.advance_positions(40 * DT_INTERNAL)
        gc, oc = assert_same_layout(sim.download(), osim.state(), f"drift {k}")
        assert np.array_equal(gc["r"], oc["r"]), f"drift {k}: positions differ"
        assert np.array_equal(gc["p"], oc["p"])
        moved += int(np.sum(before != oc["id"]))
    assert moved > 100  # the passes really reordered cells


@pytest.mark.parametrize("case", ["in_cell", "outside"])
def test_rebox_multi_hop_serial_pass(gpu, cu500, ocu500, case):
    """An atom that must move twice in one pass (arrives in a higher cell
    still outside it) takes the serial path, which must match reBoxAll.

    in_cell: every atom ends inside its cell; the neighbour lists, whose
    id -> slot map and list words the serial pass moves with the atoms, then
    give the oracle's forces, and a few steps stay with the oracle.
    outside: a second atom moves down into an already scanned cell and stays
    outside it, as reBoxAll leaves it; the reference then finds that atom's
    pairs by its cell, not by distance, which the link-cell kernels follow
    (lists off) and the lists, by design, do not."""
    print(f"[case] multi-hop {case}", flush=True)
    xyz, p, bounds = seeded_atoms(8, 3.615, 63.55, temp=300.0, disp=0.05, seed=1)
    sim, osim = make_pair(cu500, ocu500, xyz, p, bounds)
    if case == "outside":
        sim.tune(skin=0.0)
    sim.rebox()
    osim.rebox()
    sim.force()  # lists built before the multi-hop pass
    st = osim.state()
    bs = osim.info()["boxsize"]
    c = 5
    st["r"][c * 64 + 0, 0] += 2.2 * bs[0]  # two or three cells up in x
    # down in y by one cell: inside it (in_cell), or still below it (outside)
    st["r"][c * 64 + 1, 1] -= (1.5 if case == "outside" else 0.6) * bs[1]
    sim.upload(st["natoms"], st["id"], st["itype"], st["r"], st["p"])
    osim.set_state(st)
    sim.rebox()
    osim.rebox()
    gc, oc = assert_same_layout(sim.download(), osim.state(), "multi-hop")
    assert np.array_equal(gc["r"], oc["r"])
    outside = np.any((oc["r"] < 0.0) | (oc["r"] >= np.array(bs)), axis=1)
    assert outside.sum() == (1 if case == "outside" else 0)
    assert sim.list_info()["active"] == (1 if case == "in_cell" else 0)
    sim.force()
    osim.force()
    check_forces(sim.download(), osim.state(), FP64_TOL, f"multi-hop {case} force")
    for step in range(3):
        sim.nsteps(1)
        osim.nsteps(1)
        check_forces(sim.download(), osim.state(), FP64_TOL, f"multi-hop {case} step {step}")
        assert abs(sim.energy()[0] - osim.info()["e"]) <= 1e-10 * abs(osim.info()["e"])


def test_upload_rejects_bad_ids(gpu, cu500):
    """comd_upload checks the ids the neighbour lists index by: a negative id
    and an id given twice are COMD_ERR_ARG, and the context keeps working
    with a valid upload afterwards."""
    xyz, p, bounds = seeded_atoms(6, 3.615, 63.55, temp=300.0, disp=0.05, seed=2)
    sim = CA.Simulation(cu500, bounds, defgrad=0.0)
    sim.bin_atoms(xyz, p)
    st = sim.download()
    occ = np.flatnonzero((np.arange(64)[None, :] < st["natoms"][:, None]).reshape(-1))
    for what, change in (("negative", lambda ids: ids.__setitem__(occ[3], -5)),
                         ("repeated", lambda ids: ids.__setitem__(occ[7], ids[occ[2]]))):
        ids = st["id"].copy()
        change(ids)
        with pytest.raises(CA.ComdError) as ei:
            sim.upload(st["natoms"], ids, st["itype"], st["r"], st["p"])
        print(f"[bad ids] {what}: {ei.value}", flush=True)
        assert ei.value.code == 1
    sim.upload(st["natoms"], st["id"], st["itype"], st["r"], st["p"])
    sim.force()
    assert sim.list_info()["active"] == 1 and sim.list_info()["valid"] == 1


def test_overflow_reported(gpu, cu500):
    """65 atoms in one cell: the library reports COMD_ERR_OVERFLOW (the
    reference writes past the cell silently, src-flat/domains.c:269-275)."""
    bounds = (20.0, 20.0, 20.0)
    rng = np.random.default_rng(0)
    xyz = rng.uniform(0.1, 4.0, size=(65, 3))
    sim = CA.Simulation(cu500, bounds, defgrad=0.0)
    with pytest.raises(CA.ComdError) as ei:
        sim.bin_atoms(xyz, None)
    assert ei.value.code == 4


# ---------------------------------------------------------------------------
# forces in FP64

@pytest.mark.parametrize("evaluation", ["table", "cheby"])
@pytest.mark.parametrize("n,disp,temp,seed", [(6, 0.05, 300.0, 1), (10, 0.1, 300.0, 2)])
def test_eam_force_fp64(gpu, cu500, ocu500, evaluation, n, disp, temp, seed):
    print(f"[case] eam {evaluation} n={n} disp={disp} seed={seed}", flush=True)
    ocu500.set_eval(evaluation == "cheby")
    xyz, p, bounds = seeded_atoms(n, 3.615, 63.55, temp=temp, disp=disp, seed=seed)
    sim, osim = make_pair(cu500, ocu500, xyz, p, bounds, evaluation=evaluation)
    sim.rebox()
    osim.rebox()
    sim.force()
    osim.force()
    check_forces(sim.download(), osim.state(), FP64_TOL, f"eam-{evaluation}")
    pe, ke, st = sim.energy()
    oi = osim.info()
    assert abs(pe - oi["e"]) <= 1e-12 * abs(oi["e"])
    assert abs(st - oi["stress"]) <= 1e-10 * abs(oi["stress"]) + 1e-18
    assert abs(ke - osim.kinetic()) <= 1e-12 * osim.kinetic()
    ocu500.set_eval(False)


def test_eam_perfect_lattice_closed_form(gpu, cu500):
    """Perfect FCC Cu at a=3.615: every atom has E = -0.1300992269 Ha
    (-3.5401 eV, the Rose-fit cohesive energy of pots/cu500.doc) and zero
    force.  The reference prints -4163.175260709502 Ha for 20^3; its
    sequential etot sum of ~700k terms is off by 5.7e-8 Ha from the exactly
    rounded sum of the per-atom energies (-4163.175260766743), so the total is
    compared at 1e-10 relative."""
    sim, _, _ = CA.fcc_simulation(cu500, 20)
    sim.rebox()
    sim.force()
    pe, _, stress = sim.energy()
    assert abs(pe - (-4163.175260709502)) <= 1e-10 * 4163.2
    st = compact(sim.download())
    assert st["f"].shape[0] == 32000
    assert np.abs(st["f"][:, :3]).max() < 1e-12
    np.testing.assert_allclose(st["f"][:, 3], -0.13009922689896, rtol=1e-12)
    assert abs(pe / 32000 * CA.HARTREE_EV - (-3.5401)) < 1e-4
    assert abs(stress - (-0.030766284910462717)) < 1e-10 * 0.0308


def test_two_atom_fixture(gpu, cu500, ocu500):
    """data/2.inp.gz: two atoms seen through three periodic faces of a 40 A
    box (image separation 2.0712 A)."""
    bounds, xyz, p = CA.read_gz(os.path.join(GOLDEN, "data", "2.inp.gz"))
    sim, osim = make_pair(cu500, ocu500, xyz, p, bounds)
    sim.rebox()
    osim.rebox()
    for k in range(3):
        if k == 0:
            sim.force()
            osim.force()
        else:
            sim.nsteps(10)
            osim.nsteps(10)
        g, o = compact(sim.download()), compact(osim.state())
        assert np.array_equal(g["id"], o["id"])
        np.testing.assert_allclose(g["f"], o["f"], rtol=1e-10, atol=1e-16)
        pe = sim.energy()[0]
        assert abs(pe - osim.info()["e"]) <= 1e-12 * abs(osim.info()["e"])
    # trace of the reference at steps 0/10/20 (src-flat probe values)
    assert abs(pe - (-0.075811819287074477813)) < 1e-13


def lj_pot_pair():
    pot = CA.copper_lj()
    opot = O.OraclePot.lj(pot.sigma, pot.epsilon, pot.cutoff, pot.mass)
    return pot, opot


@pytest.mark.parametrize("n,temp,seed", [(5, 600.0, 4), (8, 600.0, 9)])
def test_lj_force_fp64(gpu, n, temp, seed):
    pot, opot = lj_pot_pair()
    xyz, p, bounds = seeded_atoms(n, pot.lat, pot.mass, temp=temp, disp=0.05, seed=seed)
    sim, osim = make_pair(pot, opot, xyz, p, bounds)
    sim.rebox()
    osim.rebox()
    sim.force()
    osim.force()
    check_forces(sim.download(), osim.state(), FP64_TOL, "lj", lj=True)
    pe, ke, st = sim.energy()
    oi = osim.info()
    assert abs(pe - oi["e"]) <= 1e-12 * abs(oi["e"])
    assert abs(st - oi["stress"]) <= 1e-10 * abs(oi["stress"])


def brute_neighbour_counts(xyz_slot_order, bounds, rc):
    X = np.asarray(xyz_slot_order)
    L = np.asarray(bounds)
    d = X[:, None, :] - X[None, :, :]
    d -= L * np.round(d / L)
    r2 = (d ** 2).sum(-1)
    np.fill_diagonal(r2, np.inf)
    return (r2 <= rc * rc).sum(1)


@pytest.mark.parametrize("kind,n,disp,seed", [("eam", 6, 0.05, 1), ("eam", 7, 0.1, 3), ("eam", 10, 0.1, 2),
                                              ("lj", 5, 0.05, 4), ("lj", 6, 0.05, 5), ("lj", 8, 0.05, 9)])
def test_neighbour_counts_exact(gpu, cu500, kind, n, disp, seed):
    """Per-atom number of neighbours inside the cutoff (an integer) equals a
    brute-force minimum-image count, for several cell occupancies."""
    pot = cu500 if kind == "eam" else CA.copper_lj()
    lat = 3.615 if kind == "eam" else pot.lat
    xyz, p, bounds = seeded_atoms(n, lat, pot.mass, temp=300.0, disp=disp, seed=seed)
    sim = CA.Simulation(pot, bounds, defgrad=0.0)
    assert min(sim.nbx) >= 3
    sim.bin_atoms(xyz, p)
    sim.rebox()
    sim.force()
    st = sim.download()
    nnb, _ = sim.download_aux()
    m = (np.arange(64)[None, :] < st["natoms"][:, None]).reshape(-1)
    ref = brute_neighbour_counts(xyz[st["id"][m]], bounds, pot.cutoff)
    bad = np.flatnonzero(nnb[m] != ref)
    print(f"[{kind} n={n}] atoms={m.sum()} occupancy {np.bincount(st['natoms']).nonzero()[0]} mismatches {len(bad)}",
          flush=True)
    assert len(bad) == 0
    assert sim.pair_count() == ref.sum() // 2


@pytest.mark.parametrize("n", [2, 3])
@pytest.mark.parametrize("evaluation", ["table", "cheby"])
def test_small_grids_match_reference_semantics(gpu, cu500, ocu500, n, evaluation):
    """Boxes with 1 or 2 cells per dimension: a single cell sees only itself
    (src-flat/domains.c:406-410); with 2 cells the 27-cell stencil repeats
    cells and the reference counts those pairs repeatedly (domains.c:419-426).
    Both are reproduced."""
    ocu500.set_eval(evaluation == "cheby")
    xyz, p, bounds = seeded_atoms(n, 3.615, 63.55, temp=300.0, disp=0.05, seed=7)
    sim, osim = make_pair(cu500, ocu500, xyz, p, bounds, evaluation=evaluation)
    assert max(sim.nbx) <= 2
    sim.rebox()
    osim.rebox()
    sim.force()
    osim.force()
    check_forces(sim.download(), osim.state(), FP64_TOL, f"small-grid n={n} {evaluation}")
    assert abs(sim.energy()[0] - osim.info()["e"]) <= 1e-12 * abs(osim.info()["e"])
    ocu500.set_eval(False)


# ---------------------------------------------------------------------------
# timestepping: per step from a shared state

@pytest.mark.parametrize("kind", ["eam", "lj"])
def test_steps_from_shared_state(gpu, cu500, ocu500, kind):
    if kind == "eam":
        pot, opot = cu500, ocu500
        xyz, p, bounds = seeded_atoms(8, 3.615, 63.55, temp=600.0, disp=0.1, seed=13)
    else:
        pot, opot = lj_pot_pair()
        xyz, p, bounds = seeded_atoms(6, pot.lat, pot.mass, temp=600.0, disp=0.0, seed=13)
    sim, osim = make_pair(pot, opot, xyz, p, bounds)
    sim.rebox()
    osim.rebox()
    for step in range(5):
        st = sim.download()
        osim.set_state(st)
        sim.nsteps(1)
        osim.nsteps(1)
        g, o = assert_same_layout(sim.download(), osim.state(), f"{kind} step {step}")
        np.testing.assert_allclose(g["r"], o["r"], rtol=0, atol=1e-12)
        np.testing.assert_allclose(g["p"], o["p"], rtol=1e-10, atol=1e-14 * np.abs(o["p"]).max())
        ef = force_rel_error(g["f"], o["f"])
        assert ef.max() <= FP64_TOL


@pytest.mark.parametrize("kind", ["eam", "lj"])
def test_nonperiodic_cluster(gpu, cu500, ocu500, kind):
    """The reference's non-periodic mode (PERIODIC = 0): a hot lattice block
    in a box with room around it; the stencil leaves out cells beyond the
    faces (src-flat/domains.c:318) and reBoxAll keeps atoms that cross a
    face in their edge cell (domains.c:375-378).  Forces at 1e-10 and 10
    steps, each from the oracle's state, with slot order exact."""
    print(f"[case] non-periodic {kind}", flush=True)
    if kind == "eam":
        pot, opot, lat = cu500, ocu500, 3.615
    else:
        pot, opot = lj_pot_pair()
        lat = pot.lat
    xyz, p, _ = seeded_atoms(5, lat, pot.mass, temp=3000.0, disp=0.05, seed=17)
    xyz = xyz + 2.0 * lat  # the block away from the low faces
    bounds = (9 * lat,) * 3
    sim, osim = make_pair(pot, opot, xyz, p, bounds, periodic=False)
    sim.rebox()
    osim.rebox()
    assert sim.list_info()["active"] == 0  # lists need a periodic box
    sim.force()
    osim.force()
    check_forces(sim.download(), osim.state(), FP64_TOL, f"non-periodic {kind} force", lj=kind == "lj")
    for step in range(10):
        osim.set_state(sim.download())
        sim.nsteps(1)
        osim.nsteps(1)
        g, o = assert_same_layout(sim.download(), osim.state(), f"non-periodic {kind} step {step}")
        np.testing.assert_allclose(g["r"], o["r"], rtol=0, atol=1e-12)
        assert force_rel_error(g["f"], o["f"]).max() <= FP64_TOL


def test_trajectory_energy(gpu, cu500, ocu500):
    """20 steps on both sides: the energies stay together and the DKD
    integrator conserves PE + KE for EAM (f = +grad U)."""
    xyz, p, bounds = seeded_atoms(8, 3.615, 63.55, temp=300.0, disp=0.05, seed=17)
    sim, osim = make_pair(cu500, ocu500, xyz, p, bounds)
    sim.rebox()
    osim.rebox()
    sim.force()
    pe0, ke0, _ = sim.energy()
    for _ in range(2):
        eg = sim.nsteps(10)
        eo = osim.nsteps(10)
        assert abs(eg - eo) <= 1e-9 * abs(eo)
    pe, ke, _ = sim.energy()
    assert abs((pe + ke) - (pe0 + ke0)) < 2e-4 * abs(ke0)


# ---------------------------------------------------------------------------
# mixed precision: the potential's values in FP32, its derivatives, the pair
# geometry, sums and integration in FP64, list positions in 32-bit fixed
# point.  Bar: per-atom force error <= 1e-6 against the FP64 oracle.

@pytest.mark.parametrize("evaluation,tol_rho", [("table", 1e-6), ("cheby", 5e-6)])
def test_eam_mixed_precision_error(gpu, cu500, ocu500, evaluation, tol_rho):
    ocu500.set_eval(evaluation == "cheby")
    xyz, p, bounds = seeded_atoms(10, 3.615, 63.55, temp=300.0, disp=0.05, seed=2)
    sim, osim = make_pair(cu500, ocu500, xyz, p, bounds, evaluation=evaluation, precision="mixed")
    sim.rebox()
    osim.rebox()
    sim.force()
    osim.force()
    g, o = assert_same_layout(sim.download(), osim.state(), "mixed")
    ef = force_rel_error(g["f"], o["f"])
    ee = scalar_rel_error(g["f"][:, 3], o["f"][:, 3])
    er = scalar_rel_error(g["rho"], o["rho"])
    pe = sim.energy()[0]
    print(f"[mixed {evaluation}] force max {ef.max():.3e} median {np.median(ef):.3e}; energy/atom max {ee.max():.3e}; "
          f"rho max {er.max():.3e}; total PE rel {abs(pe - osim.info()['e']) / abs(osim.info()['e']):.3e}", flush=True)
    ocu500.set_eval(False)
    assert ef.max() < 1e-6
    assert er.max() < tol_rho


def test_lj_mixed_precision_error(gpu):
    pot, opot = lj_pot_pair()
    xyz, p, bounds = seeded_atoms(6, pot.lat, pot.mass, temp=600.0, disp=0.05, seed=8)
    sim, osim = make_pair(pot, opot, xyz, p, bounds, precision="mixed")
    sim.rebox()
    osim.rebox()
    sim.force()
    osim.force()
    g, o = assert_same_layout(sim.download(), osim.state(), "lj mixed")
    ef = force_rel_error(g["f"], o["f"])
    print(f"[lj mixed] force max {ef.max():.3e} median {np.median(ef):.3e}", flush=True)
    assert ef.max() < 1e-6


# ---------------------------------------------------------------------------
# Lennard-Jones force convention (COMD_TUNE_LJ_PHYSICAL)

def test_lj_physical_sign_negates_forces_only(gpu):
    """The physical convention stores +grad U: forces are the exact negation
    of the reference convention's, energies and neighbour counts unchanged."""
    pot = CA.copper_lj()
    xyz, p, bounds = seeded_atoms(6, pot.lat, pot.mass, temp=300.0, disp=0.05, seed=21)
    a = CA.Simulation(pot, bounds, defgrad=0.0)
    b = CA.Simulation(pot, bounds, defgrad=0.0)
    b.tune(lj_physical=1)
    for s in (a, b):
        s.bin_atoms(xyz, p)
        s.rebox()
        s.force()
    ga, gb = compact(a.download()), compact(b.download())
    assert np.array_equal(ga["id"], gb["id"])
    assert np.array_equal(gb["f"][:, :3], -ga["f"][:, :3])
    assert np.array_equal(gb["f"][:, 3], ga["f"][:, 3])
    assert a.energy()[0] == b.energy()[0]


@pytest.mark.parametrize("skin", [0.4, 0.0])
def test_lj_physical_motion_conserves_energy(gpu, skin):
    """Physical convention, 100 K, 300 steps: KE + PE is conserved.  The
    cutoff, 2.35 sigma = 5.44 A, sits 0.24 A from the nearest FCC shells
    (5.20 and 5.81 A), so no pair crosses it at this temperature: an
    unshifted cutoff that pairs cross (2.5 sigma = 5.79 A is 0.02 A inside the
    fifth shell) makes the total jump by U(rc) at every crossing.  In src-flat's
    convention the same lattice flies apart (oracle CPU test)."""
    s6 = 2.315
    pot = CA.LJPotential(sigma=s6, epsilon=0.167 / CA.HARTREE_EV, cutoff=2.35 * s6, mass=63.55)
    lat = np.sqrt(2.0) * 2.0 ** (1.0 / 6.0) * s6
    xyz, p, bounds = seeded_atoms(8, lat, pot.mass, temp=100.0, disp=0.0, seed=22)
    s = CA.Simulation(pot, bounds, defgrad=0.0)
    s.tune(lj_physical=1, skin=skin)
    s.bin_atoms(xyz, p)
    s.rebox()
    s.force()
    pe0, ke0, _ = s.energy()
    e, k = [], []
    for _ in range(6):
        s.nsteps(50)
        pe, ke, _ = s.energy()
        e.append(pe + ke)
        k.append(ke)
    drift = max(abs(x - (pe0 + ke0)) for x in e) / ke0
    print(f"[lj physical skin={skin}] KE0 {ke0:.4e} KE {k} max |dE|/KE0 {drift:.2e}", flush=True)
    assert drift < 1e-2
    assert all(0.05 * ke0 < x < 2.0 * ke0 for x in k)  # bound: no runaway heating
