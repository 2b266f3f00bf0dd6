"""Observables beyond the reference's (comd_observe, comd_run_monitored):
the full virial tensor, temperature and pressure, and the on-device energy
monitor, against the oracle's tensor (orc_virial_tensor, itself pinned by
the strain derivative of the oracle's energy in tests/test_oracle.py) and
against closed forms.  The reference's own scalar "stress" stays the parity
quantity of the force passes and is checked here as part of the same
reduction."""
import numpy as np
import pytest

import comd_amd as CA
import oracle_py as O
from helpers import seeded_atoms

pytestmark = pytest.mark.gpu


def pair(pot, opot, n, lat, temp, disp, seed, evaluation="table", precision="fp64", physical=False):
    xyz, p, bounds = seeded_atoms(n, lat, pot.mass, temp=temp, disp=disp, seed=seed)
    sim = CA.Simulation(pot, bounds, defgrad=0.0, precision=precision, evaluation=evaluation)
    if physical:
        sim.tune(lj_physical=1)
    sim.bin_atoms(xyz, p)
    osim = O.OracleSim.from_atoms(opot, bounds, xyz, p)
    sim.rebox()
    osim.rebox()
    return sim, osim


@pytest.mark.parametrize("kind,evaluation", [("eam", "table"), ("eam", "cheby"), ("lj", "table"),
                                             ("lj_physical", "table")])
def test_virial_tensor_matches_oracle(gpu, cu500, ocu500, kind, evaluation):
    print(f"[case] virial {kind} {evaluation}", flush=True)
    if kind == "eam":
        pot, opot, lat = cu500, ocu500, 3.615
        ocu500.set_eval(evaluation == "cheby")
    else:
        pot = CA.copper_lj()
        opot = O.OraclePot.lj(pot.sigma, pot.epsilon, pot.cutoff, pot.mass)
        lat = pot.lat
    try:
        sim, osim = pair(pot, opot, 7, lat, 300.0, 0.1, 3, evaluation=evaluation, physical=kind == "lj_physical")
        g = sim.observe()
        osim.force()
        w, wp = osim.virial()
        oi = osim.info()
        scale = np.abs(w).max()
        print(f"[virial {kind}] W diag {np.diag(g['virial'])} max |dW| {np.abs(g['virial'] - w).max():.2e} "
              f"of {scale:.3e}; T {g['temperature']:.2f} K; P {g['pressure'] * 4359.74:.4f} GPa", flush=True)
        assert np.abs(g["virial"] - w).max() <= 1e-10 * scale
        assert np.abs(g["virial_pair"] - wp).max() <= 1e-10 * np.abs(wp).max()
        np.testing.assert_allclose(g["virial"], g["virial"].T, rtol=0, atol=1e-13 * scale)
        assert abs(g["pe"] - oi["e"]) <= 1e-12 * abs(oi["e"])
        assert abs(g["ke"] - osim.kinetic()) <= 1e-12 * osim.kinetic()
        assert abs(g["stress_ref"] - oi["stress"]) <= 1e-10 * abs(oi["stress"])
        assert g["natoms"] == oi["ntot"]
        kb = 3.166811563e-6
        assert abs(g["temperature"] - 2 * g["ke"] / (3 * g["natoms"] * kb)) <= 1e-12 * g["temperature"]
        assert abs(g["pressure"] - np.trace(g["virial"]) / (3 * g["volume"])) <= 1e-14 * abs(np.trace(w))
    finally:
        ocu500.set_eval(False)


def test_virial_mixed_precision(gpu, cu500, ocu500):
    sim, osim = pair(cu500, ocu500, 8, 3.615, 300.0, 0.05, 4, precision="mixed")
    g = sim.observe()
    osim.force()
    w, _ = osim.virial()
    err = np.abs(g["virial"] - w).max() / np.abs(w).max()
    print(f"[virial mixed] max relative error {err:.2e}", flush=True)
    assert err <= 1e-6


def test_pressure_is_volume_derivative_of_perfect_lattice(gpu, cu500):
    """Perfect FCC Cu in Chebyshev form (whose derivative series are the
    exact derivatives of its value series): the virial is isotropic and its
    pressure equals -dE/dV of the lattice energy, by central differences
    in the lattice constant."""
    def lattice(a):
        sim, _, _ = CA.fcc_simulation(cu500, 8, lat=a, evaluation="cheby")
        sim.rebox()
        return sim.observe()

    a, h = 3.615, 1e-5
    g = lattice(a)
    wp = g["virial_pair"]
    assert np.abs(wp - np.diag(np.diag(wp))).max() <= 1e-11 * abs(wp[0, 0])
    assert np.ptp(np.diag(wp)) <= 1e-11 * abs(wp[0, 0])
    ep, em = lattice(a * (1 + h))["pe"], lattice(a * (1 - h))["pe"]
    vp, vm = g["volume"] * (1 + h) ** 3, g["volume"] * (1 - h) ** 3
    p_fd = -(ep - em) / (vp - vm)
    print(f"[lattice pressure] virial {g['pressure']:.10e} -dE/dV {p_fd:.10e} Ha/A^3", flush=True)
    assert abs(g["pressure"] - p_fd) <= 1e-6 * abs(p_fd)


def test_batch_observe_equals_single(gpu, cu500):
    lj = CA.copper_lj()
    specs = [(cu500, 3.615, 6, 11), (lj, lj.lat, 6, 12), (cu500, 3.615, 7, 13)]
    reps, atoms = [], []
    for pot, lat, n, seed in specs:
        xyz, p, bounds = seeded_atoms(n, lat, pot.mass, temp=300.0, disp=0.05, seed=seed)
        reps.append(dict(pot=pot, bounds=bounds, defgrad=0.0))
        atoms.append((xyz, p, bounds, pot))
    b = CA.Batch(reps, precision="fp64")
    for r, (xyz, p, _, _) in enumerate(atoms):
        b.bin_atoms(r, xyz, p)
    b.rebox()
    for r, (xyz, p, bounds, pot) in enumerate(atoms):
        s = CA.Simulation(pot, bounds, defgrad=0.0)
        s.bin_atoms(xyz, p)
        s.rebox()
        gs, gb = s.observe(), b.observe_of(r)
        assert np.abs(gs["virial"] - gb["virial"]).max() <= 1e-12 * np.abs(gs["virial"]).max()
        assert abs(gs["pe"] - gb["pe"]) <= 1e-12 * abs(gs["pe"])


def test_energy_monitor_equals_stepwise_calls(gpu, cu500):
    """comd_run_monitored records what nTimeSteps(k) + comd_energy give one
    call at a time, bit for bit, and EAM's KE + PE stays conserved."""
    sims = []
    for _ in range(2):
        s, _, _ = CA.fcc_simulation(cu500, 8, temp=300.0, disp=0.05, seed=5)
        s.rebox()
        sims.append(s)
    pe, ke = sims[0].run_monitored(5, 10)
    s = sims[1]
    s.force()
    ref = [s.energy()[:2]]
    for _ in range(5):
        s.nsteps(10)
        ref.append(s.energy()[:2])
    ref = np.array(ref)
    np.testing.assert_array_equal(pe, ref[:, 0])
    np.testing.assert_array_equal(ke, ref[:, 1])
    e = pe + ke
    print(f"[monitor] E {e}", flush=True)
    assert np.abs(e - e[0]).max() <= 2e-4 * ke[0]
