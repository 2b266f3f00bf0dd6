"""GPU tests of batches: independent replicas of uneven size, x-stretch and
potential (EAM table, EAM Chebyshev, Lennard-Jones) in one context, where
every kernel of a step covers all replicas in one launch and steps replay a
captured HIP graph.

Each replica must evolve exactly as the same system run alone (itself
parity-tested against the oracle): same cells and slot order, forces and
trajectories equal to 1e-12.  Graph replays are checked chunk by chunk with
eager work (energies, downloads) between them.
"""
import numpy as np
import pytest

import comd_amd as CA
import oracle_py as O
from helpers import POTS, compact, force_rel_error, seeded_atoms

pytestmark = pytest.mark.gpu

LJ = CA.copper_lj()


def replica_specs(seed=0, count=6):
    rng = np.random.default_rng(seed)
    specs = []
    for k in range(count):
        kind = ("eam", "lj", "eam_cheby")[k % 3]
        n = int(rng.integers(5, 9))
        stretch = float(rng.uniform(1.0, 1.1))
        specs.append(dict(kind=kind, n=n, stretch=stretch, seed=100 + k))
    return specs


def build(spec, cu500):
    pot = LJ if spec["kind"] == "lj" else cu500
    lat = LJ.lat if spec["kind"] == "lj" else 3.615
    xyz, p, bounds = seeded_atoms(spec["n"], lat, pot.mass, temp=300.0, disp=0.05, seed=spec["seed"],
                                  defgrad=spec["stretch"])
    ev = "cheby" if spec["kind"] == "eam_cheby" else "table"
    return pot, xyz, p, bounds, ev


def make_batch(specs, cu500, precision="fp64"):
    reps, atoms = [], []
    for sp in specs:
        pot, xyz, p, bounds, ev = build(sp, cu500)
        reps.append(dict(pot=pot, bounds=bounds, defgrad=0.0, evaluation=ev))
        atoms.append((xyz, p))
    b = CA.Batch(reps, precision=precision)
    for r, (xyz, p) in enumerate(atoms):
        b.bin_atoms(r, xyz, p)
    return b


def make_singles(specs, cu500, precision="fp64"):
    out = []
    for sp in specs:
        pot, xyz, p, bounds, ev = build(sp, cu500)
        s = CA.Simulation(pot, bounds, defgrad=0.0, precision=precision, evaluation=ev)
        s.bin_atoms(xyz, p)
        out.append(s)
    return out


def compare(b, singles, what, tol=1e-12, rtol=1e-12, etol=1e-12):
    for r, s in enumerate(singles):
        g, o = b.download(r), s.download()
        assert np.array_equal(g["natoms"], o["natoms"]), f"{what} replica {r}: cell counts"
        gc, oc = compact(g), compact(o)
        assert np.array_equal(gc["id"], oc["id"]), f"{what} replica {r}: slot order"
        ef = force_rel_error(gc["f"], oc["f"]).max()
        dr = np.abs(gc["r"] - oc["r"]).max()
        dp = np.abs(gc["p"] - oc["p"]).max() / max(np.abs(oc["p"]).max(), 1e-300)
        assert ef <= tol and dr <= rtol and dp <= tol, f"{what} replica {r}: f {ef:.2e} r {dr:.2e} p {dp:.2e}"
        pe_b = b.energy_of(r)[0]
        pe_s = s.energy()[0]
        assert abs(pe_b - pe_s) <= etol * abs(pe_s), f"{what} replica {r}: energy"


def test_batch_force_equals_single_runs(gpu, cu500):
    specs = replica_specs(1)
    b = make_batch(specs, cu500)
    singles = make_singles(specs, cu500)
    for r, s in enumerate(singles):
        assert b.nbx_list[r] == s.nbx
    b.rebox()
    b.force()
    for s in singles:
        s.rebox()
        s.force()
    print(f"[batch] {len(specs)} replicas, {b.ncells} cells, grids {b.nbx_list}", flush=True)
    compare(b, singles, "force")


def test_batch_graph_steps_equal_single_runs(gpu, cu500):
    specs = replica_specs(2)
    b = make_batch(specs, cu500)
    singles = make_singles(specs, cu500)
    b.rebox()
    b.force()
    for s in singles:
        s.rebox()
        s.force()
    # graph replays in chunks, with eager calls in between
    for chunk in (3, 5, 7):
        b.steps(chunk)
        b.force()
        b.sync()
        for s in singles:
            s.steps(chunk)
            s.force()
        compare(b, singles, f"after {chunk} graph steps")
        b.energy()


def test_batch_multi_hop_in_one_replica(gpu, cu500, ocu500):
    """A rebin pass in which an atom of one replica must move twice: that
    replica's pass is replayed serially, the others' take the parallel
    path; every replica ends in the oracle's arrangement."""
    specs = [dict(kind="eam", n=6, stretch=1.0, seed=31), dict(kind="eam", n=7, stretch=1.05, seed=32),
             dict(kind="eam", n=6, stretch=1.02, seed=33)]
    b = make_batch(specs, cu500)
    b.rebox()
    oracles = []
    for r, sp in enumerate(specs):
        _, xyz, p, bounds, _ = build(sp, cu500)
        o = O.OracleSim.from_atoms(ocu500, bounds, xyz, p)
        o.rebox()
        oracles.append(o)
    # replica 1: one atom two cells up in x, rebinned in the same pass as
    # the drift of every replica's atoms
    st = b.download(1)
    bs = oracles[1].info()["boxsize"]
    st["r"][3 * 64, 0] += 2.2 * bs[0]
    b.upload(1, st["natoms"], st["id"], st["itype"], st["r"], st["p"])
    oracles[1].set_state(st)
    dt = 30 * 1.0e-15 * CA.BOHR_PER_ATU_TO_A_PER_S
    b.advance_positions(dt)
    for o in oracles:
        o.advance_positions(dt)
    for r, o in enumerate(oracles):
        g, oc = b.download(r), o.state()
        assert np.array_equal(g["natoms"], oc["natoms"]), f"replica {r}: cell counts"
        gc, occ = compact(g), compact(oc)
        assert np.array_equal(gc["id"], occ["id"]), f"replica {r}: slot order"
        assert np.array_equal(gc["r"], occ["r"]), f"replica {r}: positions"


def test_batch_replica_against_oracle(gpu, cu500, ocu500):
    """One EAM and one LJ replica of a batch, 10 steps, against the oracle."""
    specs = [dict(kind="eam", n=6, stretch=1.04, seed=7), dict(kind="lj", n=5, stretch=1.08, seed=8)]
    b = make_batch(specs, cu500)
    b.rebox()
    b.force()
    pe = b.nsteps(10)
    olj = O.OraclePot.lj(LJ.sigma, LJ.epsilon, LJ.cutoff, LJ.mass)
    total = 0.0
    for r, sp in enumerate(specs):
        pot, xyz, p, bounds, _ = build(sp, cu500)
        osim = O.OracleSim.from_atoms(ocu500 if sp["kind"] == "eam" else olj, bounds, xyz, p)
        osim.rebox()
        osim.force()
        osim.nsteps(10)
        g, o = compact(b.download(r)), O.compact(osim.state())
        assert np.array_equal(g["id"], o["id"])
        ef = force_rel_error(g["f"], o["f"]).max()
        print(f"[batch vs oracle] replica {r} ({sp['kind']}): force err {ef:.2e}", flush=True)
        assert ef <= 1e-10
        total += osim.info()["e"]
    assert abs(pe - total) <= 1e-10 * abs(total)


def test_batch_autotune_and_mixed(gpu, cu500):
    specs = replica_specs(3, count=9)
    b = make_batch(specs, cu500, precision="mixed")
    b.rebox()
    b.force()
    f0 = [b.download(r)["f"].copy() for r in range(b.nrep)]
    lists = b.list_info()["active"]  # replicas this small run on the link-cell kernels
    w, bpc, bal, ms = b.autotune()
    print(f"[batch autotune] lists {lists}: waves {w} blocks/CU {bpc} balanced {bal}: {ms:.3f} ms per force",
          flush=True)
    assert b.list_info()["active"] == lists
    b.force()
    for r in range(b.nrep):
        f = b.download(r)["f"]
        m = np.abs(f0[r][:, :3]).max()
        # another list group shape changes the mixed passes' fixed-point frames
        assert np.abs(f[:, :3] - f0[r][:, :3]).max() <= (1e-6 if lists else 1e-12) * m
    b64 = make_batch(specs, cu500, precision="fp64")
    b64.rebox()
    b64.force()
    for r in range(b.nrep):
        g, o = compact(b.download(r)), compact(b64.download(r))
        ef = force_rel_error(g["f"], o["f"]).max()
        assert ef <= 1e-4, f"mixed replica {r}: {ef:.2e}"
    b.steps(20)
    b64.steps(20)
    b.sync()
    b64.sync()
    for r in range(b.nrep):
        e32, e64 = b.energy_of(r)[0], b64.energy_of(r)[0]
        assert abs(e32 - e64) <= 1e-5 * abs(e64)


def test_batch_lj_physical_steps_equal_single_runs(gpu, cu500):
    """Physical-convention Lennard-Jones replicas beside EAM ones, 60 graph
    steps (past the point where src-flat's convention flies apart): each
    replica equals its single run."""
    specs = replica_specs(4, count=4)
    b = make_batch(specs, cu500)
    b.tune(lj_physical=1)
    singles = make_singles(specs, cu500)
    for s in singles:
        s.tune(lj_physical=1)
    b.rebox()
    b.force()
    for s in singles:
        s.rebox()
        s.force()
    ke0 = [b.energy_of(r)[1] for r in range(len(specs))]
    b.steps(60)
    b.force()
    b.sync()
    for s in singles:
        s.steps(60)
        s.force()
    compare(b, singles, "physical LJ, 60 steps", tol=1e-11)
    for r, sp in enumerate(specs):
        if sp["kind"] == "lj":
            ke = b.energy_of(r)[1]
            assert 0.2 * ke0[r] < ke < 5.0 * ke0[r], f"replica {r}: kinetic energy {ke0[r]} -> {ke} Ha"


def test_batch_neighbour_lists_equal_single_runs(gpu, cu500):
    """EAM (table) and Lennard-Jones replicas large enough for neighbour
    lists (>= 3 auxiliary cells per dimension): the batch keeps lists for
    every replica (each with its own auxiliary grid and groups, one pass per
    potential kind), and each replica's forces and graph-replayed steps equal
    the same system run alone (also on lists) and the batch on link cells."""
    rng = np.random.default_rng(7)
    specs = [dict(kind=("eam", "lj")[k % 2], n=int(rng.integers(6, 9)), stretch=float(rng.uniform(1.0, 1.1)),
                  seed=300 + k) for k in range(6)]
    for precision, tol in (("fp64", 1e-12), ("mixed", 1e-6)):
        b = make_batch(specs, cu500, precision)
        c = make_batch(specs, cu500, precision)
        c.tune(skin=0)
        singles = make_singles(specs, cu500, precision)
        b.tune(lj_physical=1)
        c.tune(lj_physical=1)
        for s in singles:
            s.tune(lj_physical=1)
        for x in [b, c] + singles:
            x.rebox()
            x.force()
        info = b.list_info()
        assert info["active"] == 1 and info["valid"] == 1, info
        assert c.list_info()["active"] == 0
        compare(b, singles, f"{precision} lists, force", tol, etol=tol)
        compare(c, singles, f"{precision} cells, force", tol, etol=tol)
        if precision == "fp64":
            for chunk in (4, 9):
                for x in [b, c] + singles:
                    x.steps(chunk)
                    x.force()
                b.sync()
                compare(b, singles, f"lists, after {chunk} graph steps", 1e-9, rtol=1e-10, etol=1e-10)
        occ = b.list_occupancy()
        print(f"[batch lists] {precision}: builds {b.list_info()['builds']}, occupancy {occ}", flush=True)
        if precision == "mixed":
            # each replica's group shape keeps the one halo capacity small
            # enough for two resident blocks per CU in both passes
            assert min(occ["blocks_per_cu"]) >= 2, occ
        for x in [b, c] + singles:
            x.close()


def test_batch_autotune_tunes_list_passes(gpu, cu500):
    """A batch whose replicas all take lists: comd_autotune times the list
    passes' shapes (the per-replica search and the five fixed group
    options, rows in LDS or not, one or two atoms per thread) and applies
    the fastest; each replica's forces after it equal its single run's."""
    lj = CA.copper_lj()
    specs = [(cu500, 3.615, 7, 1), (cu500, 3.615, 8, 2), (lj, lj.lat, 7, 3), (cu500, 3.615, 9, 4)]
    reps, atoms = [], []
    for pot, lat, n, seed in specs:
        xyz, p, bounds = seeded_atoms(n, lat, pot.mass, temp=300.0, disp=0.05, seed=seed)
        reps.append(dict(pot=pot, bounds=bounds, defgrad=0.0))
        atoms.append((xyz, p, bounds, pot))
    b = CA.Batch(reps, precision="fp64")
    b.tune(lj_physical=1)
    for r, (xyz, p, _, _) in enumerate(atoms):
        b.bin_atoms(r, xyz, p)
    b.rebox()
    assert b.list_info()["active"] == 1
    res = b.autotune_lists()
    print(f"[batch list autotune] {res} occupancy {b.list_occupancy()}", flush=True)
    assert b.list_info()["active"] == 1 and b.list_info()["valid"] == 1 and res["ms"] > 0.0
    for r, (xyz, p, bounds, pot) in enumerate(atoms):
        s = CA.Simulation(pot, bounds, defgrad=0.0)
        s.tune(lj_physical=1, skin=0.0)
        s.bin_atoms(xyz, p)
        s.rebox()
        s.force()
        g, o = compact(b.download(r)), compact(s.download())
        assert np.array_equal(g["id"], o["id"])
        assert force_rel_error(g["f"], o["f"]).max() <= 1e-12
        s.close()
