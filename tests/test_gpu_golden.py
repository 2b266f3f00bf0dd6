"""GPU parity against the compiled reference's own outputs: the HIP path,
started from the same input files, is compared with the fixtures that
tests/golden/make_golden.py wrote from the reference binaries (after the
initial force and after each further nTimeSteps(K) call).

Bars: cell counts and slot order bit-exact at every dump; per-atom forces,
energies, rho and F' within 1e-10 relative (FP64); positions and momenta
within 1e-9 relative to their scale after up to 10 steps (the force sums
are added in another order, so trajectories part in the last bits)."""
import os

import numpy as np
import pytest

import comd_amd as CA
from helpers import GOLDEN, POTS, force_rel_error, scalar_rel_error

pytestmark = pytest.mark.gpu

CASES = [
    ("two_atom_table.npz", "cu500", "table", "2.inp.gz"),
    ("two_atom_cheby.npz", "cu500", "cheby", "2.inp.gz"),
    ("cu6_table.npz", "cu500", "table", "cu6.inp.gz"),
    ("cu6_cheby.npz", "cu500", "cheby", "cu6.inp.gz"),
    ("cu6s_table.npz", "cu500", "table", "cu6s.inp.gz"),
    ("lj6.npz", "lj", "table", "lj6.inp.gz"),
    ("ag8k.npz", "ag", "table", "8k.inp.gz"),
]


def potential(name):
    if name == "lj":
        return CA.copper_lj()
    return CA.EAMPotential(name, POTS)


@pytest.mark.parametrize("fixture,potname,evaluation,inp", CASES, ids=[c[0][:-4] for c in CASES])
def test_hip_path_matches_reference_fixture(gpu, fixture, potname, evaluation, inp):
    print(f"[case] {fixture}", flush=True)
    g = np.load(os.path.join(GOLDEN, fixture))
    bounds, xyz, p = CA.read_gz(os.path.join(GOLDEN, "data", inp))
    pot = potential(potname)
    lj = potname == "lj"
    sim = CA.Simulation(pot, bounds, defgrad=0.0, evaluation=evaluation)
    assert sim.nbx == tuple(g["nbx"])
    sim.bin_atoms(xyz, p)
    sim.rebox()
    sim.force()
    for m in range(int(g["dumps"]) + 1):
        if m:
            sim.nsteps(int(g["steps"]))
        st = sim.download()
        mask = (np.arange(CA.MAXATOMS)[None, :] < st["natoms"][:, None]).reshape(-1)
        np.testing.assert_array_equal(st["natoms"], g[f"s{m}_natoms"])
        np.testing.assert_array_equal(st["id"][mask], g[f"s{m}_id"])
        f = st["f"][mask]
        ef = force_rel_error(f, g[f"s{m}_f"]).max()
        ee = scalar_rel_error(f[:, 3], g[f"s{m}_f"][:, 3]).max()
        print(f"[{fixture} dump {m}] force {ef:.2e} energy {ee:.2e}", flush=True)
        assert ef <= 1e-10 and ee <= 1e-10
        if not lj:
            assert scalar_rel_error(st["rho"][mask], g[f"s{m}_rho"]).max() <= 1e-10
            assert scalar_rel_error(st["fi"][mask], g[f"s{m}_fi"]).max() <= 1e-10
        if f"s{m}_r" in g.files:
            for k in ("r", "p"):
                ref = g[f"s{m}_{k}"]
                scale = max(np.abs(ref).max(), 1e-300)
                assert np.abs(st[k][mask] - ref).max() <= 1e-9 * scale, (m, k)
        pe = sim.energy()[0]
        ref_e = float(g[f"s{m}_e"])
        assert abs(pe - ref_e) <= 1e-10 * abs(ref_e)
