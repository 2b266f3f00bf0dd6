"""Parity at the benchmark configurations' full sizes.

The other GPU test files compare the HIP path with the oracle on a few
thousand atoms.  Here the same bars are applied at the shapes the
benchmarks run: the neighbour-list passes then run with their real group
counts, halo capacities and list capacity.

* 256 000 Cu atoms (FCC 40^3, 0.05 A displacements, 300 K): FP64 table and
  Chebyshev forces against the oracle at 1e-10 per atom, and 10 steps of the
  table path with the slot order exact afterwards.
* 2 048 000 Cu atoms (FCC 80^3): FP64 forces at 1e-10 and mixed-precision
  forces at 1e-6 against the oracle.
* 2 048 000 Cu atoms, 1000 steps from the same seeded 300 K lattice in FP64
  and in mixed precision: the total energy of the two runs, and its drift.
* 32 000-atom Lennard-Jones copper (FCC 20^3, 600 K), stepped one step at a
  time from the oracle's state: for the first 40 steps in src-flat's force
  convention (after about 50 steps the reference's lattice flies apart and
  its rebinning aborts, src-flat/domains.c:229), and for 100 steps in the
  physical convention.
* the 64-replica batch (about 1.2 M atoms, half Lennard-Jones, half EAM,
  mixed precision): sampled replicas against single contexts of their own,
  after binning and after 50 graph-replayed steps.
"""
import numpy as np
import pytest

import comd_amd as CA
import oracle_py as O
from helpers import DT_INTERNAL, compact, force_rel_error, scalar_rel_error, seeded_atoms

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-10
MIXED_TOL = 1e-6


def make_pair(pot, opot, xyz, p, bounds, precision="fp64", evaluation="table"):
    sim = CA.Simulation(pot, bounds, defgrad=0.0, precision=precision, evaluation=evaluation)
    sim.bin_atoms(xyz, p, itype=0)
    osim = O.OracleSim.from_atoms(opot, bounds, xyz, p, itype=0)
    assert osim.info()["err"] == 0
    assert tuple(osim.info()["nbx"]) == sim.nbx
    return sim, osim


def layout(g, o, what):
    assert np.array_equal(g["natoms"], o["natoms"]), f"{what}: cell counts differ"
    gc, oc = compact(g), compact(o)
    assert np.array_equal(gc["id"], oc["id"]), f"{what}: slot order differs"
    return gc, oc


def errors(gc, oc, lj=False):
    out = {"force": force_rel_error(gc["f"], oc["f"]).max(),
           "energy": scalar_rel_error(gc["f"][:, 3], oc["f"][:, 3]).max()}
    if not lj:
        out["rho"] = scalar_rel_error(gc["rho"], oc["rho"]).max()
        out["fi"] = scalar_rel_error(gc["fi"], oc["fi"]).max()
    return out


def report(what, natoms, err, sim=None):
    extra = ""
    if sim is not None:
        li = sim.list_info()
        active, valid, builds, kmax = li["active"], li["valid"], li["builds"], li["kmax"]
        extra = f" lists active={active} valid={valid} builds={builds} kmax={kmax}"
    print(f"[{what}] atoms={natoms} " + " ".join(f"{k} {v:.3e}" for k, v in err.items()) + extra, flush=True)


# ---------------------------------------------------------------------------
# 256 000 atoms

def test_cu40_table_fp64_force_and_steps(gpu, cu500, ocu500):
    print("[case] Cu 40^3 table FP64, force then 10 steps", flush=True)
    xyz, p, bounds = seeded_atoms(40, 3.615, 63.55, temp=300.0, disp=0.05, seed=1)
    sim, osim = make_pair(cu500, ocu500, xyz, p, bounds)
    sim.rebox()
    osim.rebox()
    sim.force()
    osim.force()
    gc, oc = layout(sim.download(), osim.state(), "cu40 force")
    err = errors(gc, oc)
    report("cu40 table force", len(gc["id"]), err, sim)
    assert sim.list_info()["active"] == 1 and sim.list_info()["valid"] == 1  # the list passes ran
    assert max(err.values()) <= FP64_TOL
    pe, ke, stress = sim.energy()
    oi = osim.info()
    assert abs(pe - oi["e"]) <= 1e-12 * abs(oi["e"])
    assert abs(stress - oi["stress"]) <= 1e-10 * abs(oi["stress"])
    assert abs(ke - osim.kinetic()) <= 1e-12 * osim.kinetic()

    eg = sim.nsteps(10)
    eo = osim.nsteps(10)
    gc, oc = layout(sim.download(), osim.state(), "cu40 after 10 steps")
    err = errors(gc, oc)
    report("cu40 table 10 steps", len(gc["id"]), err, sim)
    np.testing.assert_allclose(gc["r"], oc["r"], rtol=0, atol=1e-10)
    assert max(err.values()) <= FP64_TOL
    assert abs(eg - eo) <= 1e-11 * abs(eo)


def test_cu40_cheby_fp64_force(gpu, cu500, ocu500):
    print("[case] Cu 40^3 Chebyshev FP64 force", flush=True)
    ocu500.set_eval(True)
    try:
        xyz, p, bounds = seeded_atoms(40, 3.615, 63.55, temp=300.0, disp=0.05, seed=1)
        sim, osim = make_pair(cu500, ocu500, xyz, p, bounds, evaluation="cheby")
        sim.rebox()
        osim.rebox()
        sim.force()
        osim.force()
        gc, oc = layout(sim.download(), osim.state(), "cu40 cheby")
        err = errors(gc, oc)
        report("cu40 cheby force", len(gc["id"]), err, sim)
        assert max(err.values()) <= FP64_TOL
        assert abs(sim.energy()[0] - osim.info()["e"]) <= 1e-12 * abs(osim.info()["e"])
        sim.set_precision("mixed")
        sim.force()
        gm = compact(sim.download())
        errm = errors(gm, oc)
        report("cu40 cheby mixed force", len(gm["id"]), errm)
        assert errm["force"] <= MIXED_TOL
    finally:
        ocu500.set_eval(False)


# ---------------------------------------------------------------------------
# 2 048 000 atoms

def test_cu80_force_fp64_and_mixed(gpu, cu500, ocu500):
    print("[case] Cu 80^3 table, FP64 and mixed force", flush=True)
    xyz, p, bounds = seeded_atoms(80, 3.615, 63.55, temp=300.0, disp=0.05, seed=1)
    sim, osim = make_pair(cu500, ocu500, xyz, p, bounds)
    del xyz, p
    sim.rebox()
    osim.rebox()
    sim.force()
    osim.force()
    oi = osim.info()
    okin = osim.kinetic()
    g = sim.download()
    gc, oc = layout(g, osim.state(), "cu80 force")
    del osim  # the oracle's slot arrays (about 1.4 GB)
    err = errors(gc, oc)
    report("cu80 FP64 force", len(gc["id"]), err, sim)
    print(f"[cu80] list passes {sim.list_occupancy()}", flush=True)
    assert (sim.list_info()["active"], sim.list_info()["valid"]) == (1, 1)
    assert max(err.values()) <= FP64_TOL
    pe, ke, stress = sim.energy()
    assert abs(pe - oi["e"]) <= 1e-12 * abs(oi["e"])
    assert abs(stress - oi["stress"]) <= 1e-10 * abs(oi["stress"])
    assert abs(ke - okin) <= 1e-12 * okin
    del g, gc

    sim.set_precision("mixed")
    sim.force()
    gm = compact(sim.download())
    assert np.array_equal(gm["id"], oc["id"])
    errm = errors(gm, oc)
    report("cu80 mixed force", len(gm["id"]), errm, sim)
    assert (sim.list_info()["active"], sim.list_info()["valid"]) == (1, 1)
    assert errm["force"] <= MIXED_TOL
    assert errm["rho"] <= MIXED_TOL
    assert errm["energy"] <= MIXED_TOL
    pm = sim.energy()[0]
    assert abs(pm - oi["e"]) <= 1e-7 * abs(oi["e"])


def energy_series(precision, cu500, nblocks=20, block=50):
    """Total energy (PE + KE, hartree) of the seeded 300 K Cu 80^3 lattice at
    step 0 and after every block of steps."""
    xyz, p, bounds = seeded_atoms(80, 3.615, 63.55, temp=300.0, disp=0.0, seed=1)
    sim = CA.Simulation(cu500, bounds, defgrad=0.0, precision=precision)
    sim.bin_atoms(xyz, p)
    sim.rebox()
    sim.force()
    pe, ke, _ = sim.energy()
    e = [pe + ke]
    for _ in range(nblocks):
        pe = sim.nsteps(block)
        _, ke, _ = sim.energy()
        e.append(pe + ke)
    builds = sim.list_info()["builds"]
    sim.close()
    return np.array(e) / xyz.shape[0], builds


def test_cu80_energy_drift_1000_steps(gpu, cu500):
    """1000 steps of 1 fs from the same seeded state in FP64 and in mixed
    precision.  The total energy per atom of the two runs stays within
    1e-8 Ha of each other at every 50th step, and the mixed run's energy
    change over the run, and its spread over the last 800 steps, are no
    larger than the FP64 run's (within 2 % + 2e-9 Ha per atom)."""
    print("[case] Cu 80^3, 1000 steps FP64 then mixed", flush=True)
    e64, b64 = energy_series("fp64", cu500)
    print(f"[drift fp64] builds {b64} E/atom {e64[0]:.12f} -> {e64[-1]:.12f}", flush=True)
    emx, bmx = energy_series("mixed", cu500)
    print(f"[drift mixed] builds {bmx} E/atom {emx[0]:.12f} -> {emx[-1]:.12f}", flush=True)
    d64, dmx = e64[-1] - e64[0], emx[-1] - emx[0]
    s64, smx = np.std(e64[4:]), np.std(emx[4:])
    steps = np.arange(len(e64)) * 50.0
    k64 = np.polyfit(steps[4:], e64[4:], 1)[0]
    kmx = np.polyfit(steps[4:], emx[4:], 1)[0]
    gap = np.abs(emx - e64).max()
    print(f"[drift] change per atom fp64 {d64:.4e} mixed {dmx:.4e}; spread fp64 {s64:.3e} mixed {smx:.3e}; "
          f"slope per atom-step fp64 {k64:.3e} mixed {kmx:.3e}; max |E_mixed - E_fp64| per atom {gap:.3e}",
          flush=True)
    print("[drift] fp64 " + " ".join(f"{x:.10f}" for x in e64), flush=True)
    print("[drift] mixed " + " ".join(f"{x:.10f}" for x in emx), flush=True)
    assert gap <= 1e-8
    assert abs(dmx) <= 1.02 * abs(d64) + 2e-9
    assert smx <= 1.02 * s64 + 2e-9


# ---------------------------------------------------------------------------
# 32 000-atom Lennard-Jones copper, one step at a time from a shared state

def lj_per_step(physical, nsteps, skin):
    pot = CA.copper_lj()
    opot = O.OraclePot.lj(pot.sigma, pot.epsilon, pot.cutoff, pot.mass)
    opot.set_lj_physical(physical)
    xyz, p, bounds = seeded_atoms(20, pot.lat, pot.mass, temp=600.0, disp=0.0, seed=7)
    sim, osim = make_pair(pot, opot, xyz, p, bounds)
    sim.tune(lj_physical=1 if physical else 0, skin=skin)
    sim.rebox()
    osim.rebox()
    osim.force()
    ke0 = osim.kinetic()
    worst = {"force": 0.0, "energy": 0.0}
    for step in range(nsteps):
        st = osim.state()
        sim.upload(st["natoms"], st["id"], st["itype"], st["r"], st["p"])
        eg = sim.nsteps(1)
        eo = osim.nsteps(1)
        assert osim.info()["err"] == 0, f"oracle stopped at step {step}"
        g, o = layout(sim.download(), osim.state(), f"lj step {step}")
        pscale = max(np.abs(o["p"]).max(), 1e-300)
        # positions to 1e-10 A, or to 1e-14 relative for atoms that the
        # reference convention has flung far out (displacements of 1e8 A a
        # step, whose last bits follow the force sum's order)
        np.testing.assert_allclose(g["r"], o["r"], rtol=1e-14, atol=1e-10, err_msg=f"step {step}")
        np.testing.assert_allclose(g["p"], o["p"], rtol=0, atol=1e-10 * pscale, err_msg=f"step {step}")
        err = errors(g, o, lj=True)
        for k in worst:
            worst[k] = max(worst[k], err[k])
        assert err["force"] <= FP64_TOL, f"step {step}: force error {err['force']:.3e}"
        assert abs(eg - eo) <= 1e-11 * abs(eo)
    return worst, ke0, osim.kinetic()


def test_lj20_reference_convention_per_step(gpu):
    """src-flat's convention (f = -grad U, subtracted by the integrator):
    KE grows without bound, and by step ~50 the reference's rebinning aborts
    on a full cell.  The first 40 steps are compared one by one, with the
    link-cell kernels (lists find pairs by distance, the reference by cell
    index; the two agree while every atom sits in its cell, which the
    runaway atoms of later steps do not)."""
    print("[case] LJ 20^3 600 K, src-flat convention, 40 steps", flush=True)
    worst, ke0, ke = lj_per_step(False, 40, 0.0)
    print(f"[lj20 reference convention] worst {worst} KE {ke0:.4f} -> {ke:.4f} Ha", flush=True)
    assert ke > ke0  # the heating of this convention (physical motion would cool a 600 K lattice)


def test_lj20_physical_convention_100_steps(gpu):
    print("[case] LJ 20^3 600 K, physical convention, 100 steps", flush=True)
    worst, ke0, ke = lj_per_step(True, 100, 0.4)
    print(f"[lj20 physical] worst {worst} KE {ke0:.4f} -> {ke:.4f} Ha", flush=True)
    assert 0.3 * ke0 < ke < 1.2 * ke0


# ---------------------------------------------------------------------------
# the 64-replica batch at its size


def test_batch64_replicas_at_size(gpu, cu500):
    """The batch configuration at its size: 64 replicas, FCC n^3 with n
    uniform in 8..24, x-stretch uniform in [1.00, 1.10], even replicas
    Lennard-Jones copper (physical force convention), odd ones EAM Cu,
    300 K, mixed precision with neighbour lists, in one context (about
    1.4 M atoms; the lattices displaced by 0.05 A so that the forces
    compared at the start are not the perfect lattice's zeros).  Sampled replicas -- the smallest, the largest, the most
    stretched of each kind -- against a single context of their own: slot
    order exact and forces within 1e-6 after binning (the single one in
    FP64 over link cells), and after 50 steps replayed from the batch's
    step graph, against the single context stepped alike in mixed
    precision (the same trajectory up to the two contexts' roundings: their
    group frames quantise positions differently): slot order exact,
    positions within 5e-8 A, forces and energies within 2e-6 (measured:
    4e-9 A, 4e-7, 2e-7)."""
    rng = np.random.default_rng(7)
    lj = CA.copper_lj()
    reps, atoms = [], []
    for r in range(64):
        n = int(rng.integers(8, 25))
        stretch = float(rng.uniform(1.0, 1.10))
        pot, lat = (lj, lj.lat) if r % 2 == 0 else (cu500, 3.615)
        xyz = CA.fcc_positions(n, n, n, lat, stretch)
        xyz, p = CA.seed_state(xyz, pot.mass, 300.0, 0.05, 7000 + r)
        reps.append(dict(pot=pot, bounds=(n * lat * stretch, n * lat, n * lat), defgrad=0.0))
        atoms.append((xyz, p, n, stretch))
    total = sum(a[0].shape[0] for a in atoms)
    print(f"[case] batch of 64 replicas, {total} atoms", flush=True)
    b = CA.Batch(reps, precision="mixed")
    b.tune(lj_physical=1)
    for r, (xyz, p, _, _) in enumerate(atoms):
        b.bin_atoms(r, xyz, p)
    b.rebox()
    b.force()
    assert b.list_info()["active"] == 1

    sizes = [a[2] for a in atoms]
    pick = set()
    for kind in (0, 1):
        idx = [r for r in range(64) if r % 2 == kind]
        pick |= {min(idx, key=lambda r: sizes[r]), max(idx, key=lambda r: sizes[r]),
                 max(idx, key=lambda r: atoms[r][3])}

    def single(r, precision, skin):
        s = CA.Simulation(reps[r]["pot"], reps[r]["bounds"], defgrad=0.0, precision=precision)
        s.tune(skin=skin)
        if r % 2 == 0:
            s.tune(lj_physical=1)
        s.bin_atoms(atoms[r][0], atoms[r][1])
        s.rebox()
        s.force()
        return s

    for r in sorted(pick):
        s = single(r, "fp64", 0.0)
        gb, gs = compact(b.download(r)), compact(s.download())
        assert np.array_equal(gb["id"], gs["id"]), r
        err = force_rel_error(gb["f"], gs["f"]).max()
        print(f"[batch64] replica {r} n={sizes[r]} {'lj' if r % 2 == 0 else 'eam'}: force vs FP64 cells {err:.2e}",
              flush=True)
        assert err <= MIXED_TOL, r
        s.close()

    singles = {r: single(r, "mixed", 0.4) for r in sorted(pick)}
    b.steps(50)
    b.force()
    for r, s in singles.items():
        s.steps(50)
        s.force()
        gb, gs = compact(b.download(r)), compact(s.download())
        assert np.array_equal(gb["id"], gs["id"]), r
        dr = np.abs(gb["r"] - gs["r"]).max()
        ef = force_rel_error(gb["f"], gs["f"]).max()
        ee = scalar_rel_error(gb["f"][:, 3], gs["f"][:, 3]).max()
        print(f"[batch64] replica {r} after 50 steps: |dr| {dr:.2e} force {ef:.2e} energy {ee:.2e}", flush=True)
        assert dr <= 5e-8 and ef <= 2e-6 and ee <= 2e-6, r
        s.close()
    b.close()
