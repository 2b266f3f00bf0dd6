"""GPU tests of the Verlet neighbour lists (COMD_TUNE_SKIN_MILLI): the list
kernels must give what the link-cell kernels give (themselves parity-tested
against the oracle and the reference's fixtures) on the same atoms: the same
neighbour counts exactly, forces, energies, rho and F' to 1e-12, and the same
trajectories, including across list rebuilds triggered on the device.

The mixed-precision list passes read positions in 32-bit periodic fixed
point (L / 2^32 per unit, 6.7e-9 A for these 29 A boxes) and F' as a float;
they are held to the mixed-precision bar, 1e-6 against the FP64 link-cell
kernels.
"""
import numpy as np
import pytest

import comd_amd as CA
from helpers import compact, force_rel_error, scalar_rel_error, seeded_atoms

pytestmark = pytest.mark.gpu


def pair_of_sims(pot, n, lat, temp, disp, seed, evaluation="table", precision="fp64"):
    xyz, p, bounds = seeded_atoms(n, lat, pot.mass, temp=temp, disp=disp, seed=seed)
    out = []
    for skin in (0.4, 0.0):
        s = CA.Simulation(pot, bounds, defgrad=0.0, evaluation=evaluation, precision=precision)
        s.tune(skin=skin)
        s.bin_atoms(xyz, p)
        s.rebox()
        out.append(s)
    return out


def assert_same(a, b, what, tol=1e-12, lj=False, stol=None):
    """Forces to tol; energies, rho and F' to stol (default tol)."""
    stol = tol if stol is None else stol
    ga, gb = a.download(), b.download()
    assert np.array_equal(ga["natoms"], gb["natoms"]), what
    ca, cb = compact(ga), compact(gb)
    assert np.array_equal(ca["id"], cb["id"]), what
    ef = force_rel_error(ca["f"], cb["f"]).max()
    ee = scalar_rel_error(ca["f"][:, 3], cb["f"][:, 3]).max()
    print(f"[{what}] lists vs cells: force {ef:.2e} energy {ee:.2e}", flush=True)
    assert ef <= tol and ee <= stol, what
    if not lj:
        assert scalar_rel_error(ca["rho"], cb["rho"]).max() <= stol
        assert scalar_rel_error(ca["fi"], cb["fi"]).max() <= stol
    na, _ = a.download_aux()
    nb, _ = b.download_aux()
    assert np.array_equal(na, nb), f"{what}: neighbour counts"
    pa, pb = a.energy(), b.energy()
    assert abs(pa[0] - pb[0]) <= stol * abs(pb[0])
    assert abs(pa[2] - pb[2]) <= max(stol, 1e-10) * max(abs(pb[2]), 1e-30)


@pytest.mark.parametrize("evaluation,precision", [("table", "fp64"), ("cheby", "fp64"), ("table", "mixed"), ("cheby", "mixed")])
def test_eam_lists_equal_cells(gpu, cu500, evaluation, precision):
    a, b = pair_of_sims(cu500, 8, 3.615, 600.0, 0.1, 3, evaluation, precision)
    if precision == "mixed":
        b.set_precision("fp64")
    a.force()
    b.force()
    info = a.list_info()
    print(f"[lists] {info}", flush=True)
    assert info["active"] == 1 and info["valid"] == 1 and info["builds"] == 1
    assert b.list_info()["active"] == 0
    # mixed Chebyshev: the value series (phi, rho) are FP32 Clenshaw sums
    # from FP32 coefficients, so energies and densities carry their rounding,
    # amplified by the series' cancellation (per atom about 5e-6 of the FP64
    # energy, the total about 2e-6); forces stay within the 1e-6 bar
    tol = 1e-12 if precision == "fp64" else 1e-6
    stol = 2e-5 if (evaluation, precision) == ("cheby", "mixed") else tol
    assert_same(a, b, f"eam {evaluation} {precision}", tol=tol, stol=stol)


@pytest.mark.parametrize("lanes", [1, 4])
@pytest.mark.parametrize("evaluation,precision", [("table", "fp64"), ("cheby", "fp64"), ("table", "mixed")])
def test_lanes_per_atom_equal_cells(gpu, cu500, lanes, evaluation, precision):
    """Small systems default to two lanes per home atom in the list passes
    (COMD_TUNE_LIST_LANES); one and four must give the link-cell results
    too (each lane takes every lanes-th chunk of the atom's list and the
    lanes add their sums)."""
    a, b = pair_of_sims(cu500, 8, 3.615, 600.0, 0.1, 5, evaluation, precision)
    a.tune(list_lanes=lanes)
    if precision == "mixed":
        b.set_precision("fp64")
    a.force()
    b.force()
    assert a.list_info()["active"] == 1 and a.list_info()["valid"] == 1
    assert a.list_lanes() == lanes and b.list_lanes() == 0
    assert_same(a, b, f"eam {evaluation} {precision}, {lanes} lanes", tol=1e-12 if precision == "fp64" else 1e-6)
    a.nsteps(10)
    b.nsteps(10)
    ca, cb = compact(a.download()), compact(b.download())
    assert np.array_equal(ca["id"], cb["id"])
    assert force_rel_error(ca["f"], cb["f"]).max() <= (1e-10 if precision == "fp64" else 5e-5)


def test_lj_lanes_equal_cells(gpu):
    lj = CA.copper_lj()
    for lanes in (1, 2, 4):
        a, b = pair_of_sims(lj, 7, lj.lat, 600.0, 0.05, 6)
        a.force()
        assert a.list_lanes() == 2  # the default below 128 atoms per CU
        a.tune(list_lanes=lanes)
        a.force()
        b.force()
        assert a.list_info()["active"] == 1 and a.list_lanes() == lanes
        assert_same(a, b, f"lj, {lanes} lanes", lj=True)


def test_mixed_lists_after_fp64_steps(gpu, cu500):
    """Steps in FP64 keep only the double list positions current; switching
    to mixed precision must refresh the fixed-point ones."""
    a, b = pair_of_sims(cu500, 7, 3.615, 1500.0, 0.05, 11)
    a.force()
    b.force()
    a.nsteps(15)
    b.nsteps(15)
    a.set_precision("mixed")
    a.force()
    b.force()
    assert a.list_info()["valid"] == 1
    assert_same(a, b, "mixed after fp64 steps", tol=1e-6)
    a.nsteps(5)
    a.set_precision("fp64")
    b.nsteps(5)
    a.force()
    b.force()
    ca, cb = compact(a.download()), compact(b.download())
    assert np.array_equal(ca["id"], cb["id"])
    assert force_rel_error(ca["f"], cb["f"]).max() <= 1e-5


def test_lj_lists_equal_cells(gpu):
    lj = CA.copper_lj()
    a, b = pair_of_sims(lj, 7, lj.lat, 600.0, 0.05, 4)
    a.force()
    b.force()
    assert a.list_info()["active"] == 1
    assert_same(a, b, "lj", lj=True)


def test_trajectory_with_rebuilds(gpu, cu500):
    """3000 K for 40 steps: atoms move well past half the skin, so the lists
    are rebuilt on the device several times; the trajectory must follow the
    link-cell one."""
    a, b = pair_of_sims(cu500, 6, 3.615, 3000.0, 0.05, 9)
    a.force()
    b.force()
    for chunk in range(4):
        pa = a.nsteps(10)
        pb = b.nsteps(10)
        assert abs(pa - pb) <= 1e-9 * abs(pb)
    info = a.list_info()
    print(f"[rebuilds] {info}", flush=True)
    assert info["valid"] == 1 and info["builds"] >= 3
    assert_same(a, b, "after 40 steps at 3000 K", tol=1e-8)


def test_upload_path_sets_up_lists(gpu, cu500):
    a, b = pair_of_sims(cu500, 6, 3.615, 300.0, 0.1, 5)
    st = b.download()
    c = CA.Simulation(cu500, a.bounds, defgrad=0.0)
    c.upload(st["natoms"], st["id"], st["itype"], st["r"], st["p"])
    c.force()
    b.force()
    assert c.list_info()["active"] == 1
    assert_same(c, b, "uploaded state")


@pytest.mark.parametrize("groups,rows,div", [(0, 1, 1), (1, 1, 2), (2, 0, 1), (3, 1, 2), (4, 1, 1), (0, 0, 2)])
@pytest.mark.parametrize("n", [5, 7])
def test_group_options_equal_cells(gpu, cu500, groups, rows, div, n):
    """Every group size (4x4x3 .. 2x2x1 auxiliary cells), table rows in LDS
    or not, one or two home atoms per thread, and boxes of 3 auxiliary cells
    per dimension (n = 5: a group's halo holds the same cell twice, at two
    images): the same pairs and forces as the link-cell kernels."""
    for prec in ("fp64", "mixed"):
        a, b = pair_of_sims(cu500, n, 3.615, 900.0, 0.1, 30 + n, "table", prec)
        a.tune(list_groups=groups, list_rows_lds=rows, list_threads_div=div)
        if prec == "mixed":
            b.set_precision("fp64")
        a.force()
        b.force()
        info = a.list_info()
        assert info["active"] == 1 and info["valid"] == 1, info
        what = f"groups {groups} rows {rows} div {div} n {n} {prec}"
        assert_same(a, b, what, tol=1e-12 if prec == "fp64" else 1e-6)
        a.nsteps(20)
        b.nsteps(20)
        assert_same(a, b, what + ", 20 steps", tol=1e-10 if prec == "fp64" else 1e-5)


def test_overflow_falls_back_to_cells(gpu, cu500):
    """A Cu 8^3 lattice whose slab x < 18 A is squeezed 2.25-fold: the lists
    are set up for the mean density, but the fullest 7.2 A auxiliary cell
    gets 72 atoms, more than the 64 one holds (the fullest 5.8 A link cell
    56).  The build marks the lists invalid and the link-cell passes,
    launched beside the list passes, give the forces -- first with the small
    grid they get while the lists were last seen valid, then with the full
    grid once the host has seen them invalid."""
    xyz, p, bounds = seeded_atoms(8, 3.615, cu500.mass, temp=0.0, disp=0.02, seed=41)
    slab = xyz[:, 0] < 18.0
    xyz[slab, 0] /= 2.25
    out = []
    for skin in (0.4, 0.0):
        s = CA.Simulation(cu500, bounds, defgrad=0.0)
        s.tune(skin=skin)
        s.bin_atoms(xyz, p)
        s.rebox()
        out.append(s)
    a, b = out
    a.force()
    b.force()
    info = a.list_info()
    print(f"[overflow] {info}", flush=True)
    assert info["active"] == 1 and info["valid"] == 0
    assert_same(a, b, "overflow, first force")
    a.nsteps(2)
    b.nsteps(2)
    assert_same(a, b, "overflow, after steps", tol=1e-10)


def test_force_after_replayed_step_graph(gpu, cu500):
    """A force call after steps replayed from an already captured step graph
    (the second steps() call reuses the graph of the first) sees the atoms'
    current positions: each step ends with a drift that leaves the list
    positions stale, whether the step ran eagerly, was captured, or was
    replayed."""
    a, b = pair_of_sims(cu500, 8, 3.615, 600.0, 0.05, 77)
    for chunk in (4, 9, 9):
        for s in (a, b):
            s.steps(chunk)
            s.force()
        assert a.list_info()["active"] == 1
        assert_same(a, b, f"after a {chunk}-step graph chunk", tol=1e-10)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_autotune_tunes_list_passes(gpu, cu500, precision):
    """comd_autotune on a context with lists times the list passes' shapes:
    the choice is applied (lists valid, occupancy reported for it), the
    forces of the state stay those of the link-cell kernels, and steps
    after it follow them."""
    a, b = pair_of_sims(cu500, 12, 3.615, 300.0, 0.05, 91, "table", precision)
    if precision == "mixed":
        b.set_precision("fp64")
    a.force()
    b.force()
    before = a.download()["f"].copy()
    res = a.autotune_lists()
    occ = a.list_occupancy()
    print(f"[autotune lists {precision}] {res} occupancy {occ}", flush=True)
    assert res["ms"] > 0.0 and res["groups"] in range(-1, 5) and res["threads_div"] in (1, 2)
    info = a.list_info()
    assert info["active"] == 1 and info["valid"] == 1
    f = a.download()["f"]
    m = np.abs(before[:, :3]).max()
    assert np.abs(f[:, :3] - before[:, :3]).max() <= (1e-12 if precision == "fp64" else 1e-6) * m
    w, bpc, bal, ms = a.autotune()  # with lists: the same tuner
    assert ms > 0.0
    tol = 1e-12 if precision == "fp64" else 1e-6
    assert_same(a, b, f"autotuned {precision}", tol=tol)
    a.nsteps(10)
    b.nsteps(10)
    assert_same(a, b, f"autotuned {precision}, 10 steps", tol=1e-10 if precision == "fp64" else 1e-5)
