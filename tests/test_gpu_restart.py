"""Restart from a snapshot (comd_snapshot_save / comd_snapshot_load): N
steps run straight through and the same N steps split by a save, a new
context and a load end in the same state, bit for bit -- slot order,
positions, momenta, forces and energies -- with the neighbour lists
rebuilt on both sides of the save."""
import numpy as np
import pytest

import comd_amd as CA
from helpers import seeded_atoms

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision,lanes", [("fp64", None), ("mixed", None), ("fp64", 4)])
def test_restart_continues_bit_for_bit(gpu, cu500, tmp_path, precision, lanes):
    """lanes: the saving context's lanes per atom in the list passes (the
    order of their sums), which the snapshot carries to a context left at
    its default."""
    xyz, p, bounds = seeded_atoms(8, 3.615, 63.55, temp=3000.0, disp=0.05, seed=61)

    def fresh():
        s = CA.Simulation(cu500, bounds, defgrad=0.0, precision=precision)
        s.bin_atoms(xyz, p)
        s.rebox()
        return s

    a = fresh()
    if lanes:
        a.tune(list_lanes=lanes)
    a.nsteps(20)
    builds_before = a.list_info()["builds"]
    path = tmp_path / "state.snap"
    a.snapshot_save(path)
    a.nsteps(30)
    ga = a.download()
    ea = a.energy()
    builds_after = a.list_info()["builds"]
    print(f"[restart {precision}] list builds {builds_before} at the save, {builds_after} at the end", flush=True)
    assert builds_after > builds_before  # the lists were rebuilt after the save

    b = CA.Simulation(cu500, bounds, defgrad=0.0, precision=precision)
    b.snapshot_load(path)
    b.nsteps(30)
    gb = b.download()
    for k in ("natoms", "id", "itype", "r", "p", "f", "rho", "fi"):
        assert np.array_equal(ga[k], gb[k]), k
    assert b.energy() == ea
