"""Centrosymmetry parameter on the GPU (comd_centrosymmetry) against the
oracle's restatement of src-viz's computeCentrosymmetry
(src-viz/AtomVisualize.c:399-673; the oracle is pinned by a brute-force
search in tests/test_oracle.py): the same neighbour search and float sums,
so every atom's value is equal, bit for bit."""
import numpy as np
import pytest

import comd_amd as CA
import oracle_py as O
from helpers import seeded_atoms

pytestmark = pytest.mark.gpu


def both(pot, opot, xyz, p, bounds):
    sim = CA.Simulation(pot, bounds, defgrad=0.0)
    sim.bin_atoms(xyz, p)
    osim = O.OracleSim.from_atoms(opot, bounds, xyz, p)
    g, o = sim.centrosymmetry(), osim.centrosymmetry()
    st = osim.state()
    m = (np.arange(64)[None, :] < st["natoms"][:, None]).reshape(-1)
    assert np.array_equal(sim.download()["id"][m], st["id"][m])
    return g[m], o[m]


@pytest.mark.parametrize("case", ["perfect", "thermal", "vacancy", "lj", "cu20", "sparse"])
def test_centrosymmetry_equals_oracle(gpu, cu500, ocu500, case):
    print(f"[case] centrosymmetry {case}", flush=True)
    if case == "sparse":  # fewer than 12 neighbours in reach: 1000 for every atom
        pot, opot = cu500, ocu500
        xyz = np.array([[1.0, 1.0, 1.0], [3.5, 1.0, 1.0], [1.0, 3.5, 1.0], [9.0, 9.0, 9.0], [15.0, 4.0, 12.0]])
        p, bounds = np.zeros_like(xyz), (20.0, 20.0, 20.0)
    elif case == "lj":
        pot = CA.copper_lj()
        opot = O.OraclePot.lj(pot.sigma, pot.epsilon, pot.cutoff, pot.mass)
        xyz, p, bounds = seeded_atoms(6, pot.lat, pot.mass, temp=600.0, disp=0.1, seed=5)
    else:
        pot, opot = cu500, ocu500
        n = 20 if case == "cu20" else 6
        disp = 0.0 if case == "perfect" else 0.1
        xyz, p, bounds = seeded_atoms(n, 3.615, 63.55, temp=300.0, disp=disp, seed=3)
        if case == "vacancy":
            xyz, p = np.delete(xyz, 100, axis=0), np.delete(p, 100, axis=0)
    g, o = both(pot, opot, xyz, p, bounds)
    print(f"[centro {case}] atoms {len(g)} range {g.min():.4g} .. {g.max():.4g}", flush=True)
    assert np.array_equal(g, o)
    if case == "perfect":
        assert np.abs(g).max() < 1e-10
    if case == "sparse":
        assert np.array_equal(g, np.full(5, 1000.0))


def test_batch_centrosymmetry_equals_single(gpu, cu500):
    """Replicas of different sizes in one context: each one's values equal
    its single simulation's (replica-local slot layout)."""
    lj = CA.copper_lj()
    specs = [(cu500, 3.615, 6, 21), (lj, lj.lat, 5, 22), (cu500, 3.615, 7, 23)]
    reps, atoms = [], []
    for pot, lat, n, seed in specs:
        xyz, p, bounds = seeded_atoms(n, lat, pot.mass, temp=300.0, disp=0.08, seed=seed)
        reps.append(dict(pot=pot, bounds=bounds, defgrad=0.0))
        atoms.append((xyz, p, bounds, pot))
    b = CA.Batch(reps, precision="fp64")
    for r, (xyz, p, _, _) in enumerate(atoms):
        b.bin_atoms(r, xyz, p)
    for r, (xyz, p, bounds, pot) in enumerate(atoms):
        print(f"[case] batch centrosymmetry replica {r}", flush=True)
        s = CA.Simulation(pot, bounds, defgrad=0.0)
        s.bin_atoms(xyz, p)
        gs, gb = s.centrosymmetry(), b.centrosymmetry_of(r)
        assert gs.shape == gb.shape
        assert np.array_equal(gs, gb)
        assert np.array_equal(s.download()["id"], b.download(r)["id"])
        s.close()
