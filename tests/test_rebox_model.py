"""CPU model of the GPU's parallel reBoxAll emulation (comd_amd/csrc/bin_rebox.hip),
checked against the oracle's serial reBoxAll (src-flat/domains.c:347-389),
and of the serial kernel (k_rebox_serial) with the id -> slot map it keeps
for the neighbour lists.

The model computes what the kernels compute -- per-atom destinations, each
touched cell's own swap-with-last scan with n_low placeholders, arrivals
ordered by (source cell, departure rank) -- on random states with many movers,
and must give the identical slot layout whenever no arrival into a higher
cell would move again (the case the GPU hands to its serial kernel).
"""
import numpy as np
import pytest

import oracle_py as O
from helpers import POTS

M = 64


def cell_xyz(c, nbx):
    return c % nbx[0], (c // nbx[0]) % nbx[1], (c // nbx[0] // nbx[1]) % nbx[2]


def wrap(ix, nbx, periodic):
    if not periodic:
        if any(ix[k] < 0 or ix[k] >= nbx[k] for k in range(3)):
            return -1
    ix = [ix[k] + nbx[k] if ix[k] < 0 else (ix[k] - nbx[k] if ix[k] >= nbx[k] else ix[k]) for k in range(3)]
    return ix[0] + ix[1] * nbx[0] + ix[2] * nbx[0] * nbx[1]


def destination(r, c, nbx, bs, periodic):
    ix = list(cell_xyz(c, nbx))
    nw, rn = list(ix), list(r)
    for k in range(3):
        if r[k] < 0.0:
            nw[k] = ix[k] - 1
            rn[k] = r[k] + bs[k]
        elif r[k] >= bs[k]:
            nw[k] = ix[k] + 1
            rn[k] = r[k] - bs[k]
    j = wrap(nw, nbx, periodic)
    return (-1 if j < 0 or j == c else j), rn


def model_rebox(st, nbx, bs, periodic=True):
    """Returns the new state, or None when the pass needs the serial path."""
    natoms = st["natoms"].copy()
    ncells = natoms.size
    ids, r, p = st["id"].copy(), st["r"].copy(), st["p"].copy()
    dest = {}
    movers = []  # (src cell, slot index, rnew)
    for c in range(ncells):
        for i in range(natoms[c]):
            s = c * M + i
            d, rn = destination(r[s], c, nbx, bs, periodic)
            if d >= 0:
                if d > c:
                    d2, _ = destination(rn, d, nbx, bs, periodic)
                    if d2 >= 0 and d2 != d:
                        return None
                dest[s] = d
                movers.append((c, i, rn))
    arrivals = {}
    for m, (c, i, rn) in enumerate(movers):
        arrivals.setdefault(dest[c * M + i], []).append(m)
    touched = set(dest[c * M + i] for c, i, _ in movers) | set(c for c, _, _ in movers)
    rank = {}
    perm = {}
    for c in touched:
        nlow = sum(1 for m in arrivals.get(c, []) if movers[m][0] < c)
        lst = list(range(natoms[c])) + [M + k for k in range(nlow)]
        k, rk = 0, 0
        while k < len(lst):
            e = lst[k]
            if e < M and (c * M + e) in dest:
                rank[c * M + e] = rk
                rk += 1
                lst[k] = lst[-1]
                lst.pop()
            else:
                k += 1
        perm[c] = lst
    new = dict(natoms=natoms.copy(), id=ids.copy(), r=r.copy(), p=p.copy(), itype=st["itype"].copy())
    for c in touched:
        arr = sorted(arrivals.get(c, []), key=lambda m: (movers[m][0], rank[movers[m][0] * M + movers[m][1]]))
        low = [m for m in arr if movers[m][0] < c]
        high = [m for m in arr if movers[m][0] > c]
        out = []
        for e in perm[c]:
            if e < M:
                s = c * M + e
                out.append((ids[s], r[s], p[s]))
            else:
                m = low[e - M]
                sc, si, rn = movers[m]
                out.append((ids[sc * M + si], rn, p[sc * M + si]))
        for m in high:
            sc, si, rn = movers[m]
            out.append((ids[sc * M + si], rn, p[sc * M + si]))
        assert len(out) < M
        for k, (a, rr, pp) in enumerate(out):
            new["id"][c * M + k] = a
            new["r"][c * M + k] = rr
            new["p"][c * M + k] = pp
        new["natoms"][c] = len(out)
    return new


@pytest.mark.parametrize("seed", range(6))
def test_model_matches_serial_rebox(seed):
    rng = np.random.default_rng(seed)
    pot = O.OraclePot.eam(POTS, "cu500")
    n = int(rng.integers(3, 6))
    lat = 3.615
    L = n * lat
    nat = 4 * n ** 3
    xyz = rng.uniform(0, L, size=(nat, 3))
    sim = O.OracleSim.from_atoms(pot, (L, L, L), xyz)
    info = sim.info()
    nbx, bs = info["nbx"], info["boxsize"]
    st = sim.state()
    # push a random 30% of the atoms by up to 0.6 cells in each dimension
    occ = np.flatnonzero((np.arange(M)[None, :] < st["natoms"][:, None]).reshape(-1))
    pick = rng.choice(occ, size=len(occ) * 3 // 10, replace=False)
    st["r"][pick] += rng.uniform(-0.6, 0.6, size=(len(pick), 3)) * bs
    sim.set_state(st)
    expect = model_rebox(st, nbx, bs)
    sim.rebox()
    got = sim.state()
    if expect is None:
        pytest.skip("pass needs the serial path")
    assert np.array_equal(got["natoms"], expect["natoms"])
    mask = (np.arange(M)[None, :] < got["natoms"][:, None]).reshape(-1)
    assert np.array_equal(got["id"][mask], expect["id"][mask])
    assert np.array_equal(got["r"][mask], expect["r"][mask])
    assert np.array_equal(got["p"][mask], expect["p"][mask])


def test_model_flags_second_hop():
    """An atom two cells up in x arrives in a higher cell still outside it."""
    pot = O.OraclePot.eam(POTS, "cu500")
    L = 5 * 3.615
    xyz = np.array([[1.0, 1.0, 1.0], [2.0, 2.0, 2.0]])
    sim = O.OracleSim.from_atoms(pot, (L, L, L), xyz)
    info = sim.info()
    st = sim.state()
    st["r"][0, 0] += 2.3 * info["boxsize"][0]
    assert model_rebox(st, info["nbx"], info["boxsize"]) is None


def model_serial(st, nbx, bs, fill_last_onto_itself=False, periodic=True):
    """The serial kernel's in-place replay of reBoxAll, with the id -> slot
    map of the neighbour lists updated by every copy (serial_copy).  The
    cell's last atom fills the hole of a departed one unless the departed
    atom was the last; fill_last_onto_itself=True makes that copy anyway
    (a self-copy), which points the departed atom's map entry back at the
    slot it left."""
    natoms = st["natoms"].copy()
    ids, r = st["id"].copy(), st["r"].copy()
    slot_of = {int(ids[c * M + i]): c * M + i for c in range(natoms.size) for i in range(natoms[c])}

    def copy(a, b):
        slot_of[int(ids[a])] = b
        ids[b] = ids[a]
        r[b] = r[a]

    for c in range(natoms.size):
        i = 0
        while i < natoms[c]:
            o = c * M + i
            j, rn = destination(r[o], c, nbx, bs, periodic)
            if j < 0:
                i += 1
                continue
            r[o] = rn
            copy(o, j * M + natoms[j])
            natoms[j] += 1
            natoms[c] -= 1
            ni = natoms[c]
            if ni > i or (fill_last_onto_itself and ni > 0):
                copy(c * M + ni, o)
    return natoms, ids, r, slot_of


@pytest.mark.parametrize("seed", range(4))
def test_serial_model_keeps_id_slot_map(seed):
    """Multi-hop passes on random states, with the last atom of some cells
    among the movers: the serial model's layout equals the oracle's reBoxAll
    and its map sends every id to the slot that holds it; the self-copy rule
    breaks the map in the same passes."""
    rng = np.random.default_rng(100 + seed)
    pot = O.OraclePot.eam(POTS, "cu500")
    n = 5
    L = n * 3.615
    xyz = rng.uniform(0, L, size=(4 * n ** 3, 3))
    sim = O.OracleSim.from_atoms(pot, (L, L, L), xyz)
    info = sim.info()
    nbx, bs = info["nbx"], info["boxsize"]
    st = sim.state()
    last = [c * M + st["natoms"][c] - 1 for c in range(st["natoms"].size) if st["natoms"][c] > 0]
    pick = rng.choice(last, size=len(last) // 3, replace=False)
    st["r"][pick, 0] += 2.3 * bs[0]  # two cells up: a multi-hop pass
    sim.set_state(st)
    assert model_rebox(st, nbx, bs) is None
    natoms, ids, r, slot_of = model_serial(st, nbx, bs)
    sim.rebox()
    got = sim.state()
    assert np.array_equal(got["natoms"], natoms)
    mask = (np.arange(M)[None, :] < natoms[:, None]).reshape(-1)
    assert np.array_equal(got["id"][mask], ids[mask])
    assert np.array_equal(got["r"][mask], r[mask])
    assert all(slot_of[int(ids[s])] == s for s in np.flatnonzero(mask))
    _, ids2, _, slot2 = model_serial(st, nbx, bs, fill_last_onto_itself=True)
    assert np.array_equal(ids2[mask], ids[mask])
    assert any(slot2[int(ids2[s])] != s for s in np.flatnonzero(mask))
