"""Shared helpers of the parity tests: seeded inputs and error measures."""
import os

import numpy as np

import comd_amd as CA

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
POTS = os.path.join(ROOT, "comd_amd", "pots")
GOLDEN = os.path.join(ROOT, "tests", "golden")
DT_INTERNAL = 1.0e-15 * CA.BOHR_PER_ATU_TO_A_PER_S


def seeded_atoms(n, lat, mass, temp=300.0, disp=0.05, seed=1, defgrad=1.0):
    """FCC n^3 in create_fcc_lattice order with seeded displacements and
    Maxwell-Boltzmann momenta; bounds stretched in x by defgrad."""
    xyz = CA.fcc_positions(n, n, n, lat, defgrad)
    p = np.zeros_like(xyz)
    if temp > 0 or disp > 0:
        xyz, p = CA.seed_state(xyz, mass, temp, disp, seed)
    bounds = (n * lat * defgrad, n * lat, n * lat)
    return xyz, p, bounds


def occupied_mask(natoms):
    return (np.arange(CA.MAXATOMS)[None, :] < np.asarray(natoms)[:, None]).reshape(-1)


def compact(st):
    m = occupied_mask(st["natoms"])
    return {k: (v[m] if k != "natoms" else v) for k, v in st.items()}


def force_rel_error(f_test, f_ref):
    """Per-atom |df| / max(|f_ref|, rms|f_ref|): relative to each atom's own
    force, floored at the RMS force for atoms whose force nearly cancels."""
    d = np.linalg.norm(f_test[:, :3] - f_ref[:, :3], axis=1)
    mag = np.linalg.norm(f_ref[:, :3], axis=1)
    rms = np.sqrt(np.mean(mag ** 2)) if mag.size else 1.0
    return d / np.maximum(mag, max(rms, 1e-300))


def scalar_rel_error(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    scale = np.maximum(np.abs(b), np.sqrt(np.mean(b ** 2)) if b.size else 1.0)
    return np.abs(a - b) / np.maximum(scale, 1e-300)
