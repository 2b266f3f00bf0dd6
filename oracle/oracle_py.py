"""ctypes view of the CPU oracle (oracle/liboracle.so) and a reader of the
dumps written by the compiled reference (oracle/_ref/ref_*).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and
by bench.py's cpu_baseline leg, always as the checker, never as the thing
measured or shipped.
"""
from __future__ import annotations

import ctypes as C
import os
import struct
import subprocess

import numpy as np

ORACLE_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(ORACLE_DIR, "liboracle.so")
REF_DIR = os.path.join(ORACLE_DIR, "_ref")
MAXATOMS = 64

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_LIB = None


def build():
    subprocess.run(["make", "-s", "-C", ORACLE_DIR, "oracle"], check=True)


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        L.orc_pot_new.restype = vp
        L.orc_pot_delete.argtypes = [vp]
        L.orc_pot_read_eam.argtypes = [vp, C.c_char_p, C.c_char_p]
        L.orc_pot_set_lj.argtypes = [vp, C.c_double, C.c_double, C.c_double, C.c_double]
        L.orc_pot_set_lj_sign.argtypes = [vp, C.c_double]
        L.orc_pot_fit_cheb.argtypes = [vp, C.c_int]
        L.orc_pot_set_eval.argtypes = [vp, C.c_int]
        L.orc_pot_get.argtypes = [vp, C.c_int]
        L.orc_pot_get.restype = C.c_double
        L.orc_pot_get_cheb.argtypes = [vp, _dp, _ip]
        L.orc_pot_get_table.argtypes = [vp, C.c_int, _ip, _dp, _dp]
        L.orc_sim_create.restype = vp
        L.orc_sim_create.argtypes = [vp, C.c_double * 3, C.c_double, C.c_double, C.c_int]
        L.orc_from_atoms.restype = vp
        L.orc_from_atoms.argtypes = [vp, C.c_double * 3, C.c_double, C.c_int, C.c_int, _dp, _dp, C.c_int]
        L.orc_create_fcc.restype = vp
        L.orc_create_fcc.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_int]
        L.orc_read_gz.restype = vp
        L.orc_read_gz.argtypes = [vp, C.c_char_p, C.c_double, C.c_int]
        L.orc_sim_free.argtypes = [vp]
        L.orc_put_atom.argtypes = [vp, C.c_int, C.c_int, C.c_double] + [C.c_double] * 6
        L.orc_rebox_all.argtypes = [vp]
        L.orc_neighbor_boxes.argtypes = [vp, C.c_int, _ip]
        L.orc_compute_force.argtypes = [vp]
        L.orc_n_time_steps.argtypes = [vp, C.c_int, C.c_double]
        L.orc_n_time_steps.restype = C.c_double
        L.orc_kinetic_energy.argtypes = [vp]
        L.orc_kinetic_energy.restype = C.c_double
        L.orc_sim_info.argtypes = [vp, _ip, _ip, _ip, _dp, _dp, _dp, _dp, _ip]
        L.orc_sim_get.argtypes = [vp, _ip, _ip, _ip, _dp, _dp, _dp, _dp, _dp]
        L.orc_sim_set_rp.argtypes = [vp, _dp, _dp]
        L.orc_table_interp.argtypes = [vp, C.c_double, _dp, _dp]
        L.orc_advance_positions.argtypes = [vp, C.c_double]
        L.orc_advance_velocity.argtypes = [vp, C.c_double]
        L.orc_sim_set_state.argtypes = [vp, _ip, _ip, _ip, _dp, _dp]
        L.orc_eam_force.argtypes = [vp]
        L.orc_lj_force.argtypes = [vp]
        L.orc_pot_eval.argtypes = [vp, C.c_int, C.c_double, _dp, _dp]
        L.orc_pot_set_cheb3.argtypes = [vp, _dp, C.c_int]
        L.orc_virial_tensor.argtypes = [vp, _dp]
        L.orc_centrosymmetry.argtypes = [vp, _dp]
        L.orc_neighbor_boxes.restype = C.c_int
        _LIB = L
    return _LIB


def _p(a, ct=C.c_double):
    return a.ctypes.data_as(C.POINTER(ct)) if a is not None else None


class OraclePot:
    def __init__(self, handle):
        self.h = handle

    @classmethod
    def eam(cls, directory, name, ncheb=32):
        L = lib()
        h = L.orc_pot_new()
        rc = L.orc_pot_read_eam(h, directory.encode(), name.encode())
        if rc:
            L.orc_pot_delete(h)
            raise RuntimeError(f"oracle could not read {directory}/{name}: {rc}")
        L.orc_pot_fit_cheb(h, ncheb)
        return cls(h)

    @classmethod
    def lj(cls, sigma, epsilon, cutoff, mass):
        L = lib()
        h = L.orc_pot_new()
        L.orc_pot_set_lj(h, sigma, epsilon, cutoff, mass)
        return cls(h)

    def set_eval(self, cheby):
        lib().orc_pot_set_eval(self.h, 1 if cheby else 0)

    def set_lj_physical(self, on):
        """LJ forces stored as +grad U (the EAM convention) instead of the
        reference's -grad U; the integrator then gives physical motion."""
        lib().orc_pot_set_lj_sign(self.h, -1.0 if on else 1.0)

    @property
    def cutoff(self):
        return lib().orc_pot_get(self.h, 0)

    @property
    def mass(self):
        return lib().orc_pot_get(self.h, 1)

    @property
    def lat(self):
        return lib().orc_pot_get(self.h, 2)

    def cheb(self):
        out = np.zeros(6 * 64)
        n = C.c_int()
        lib().orc_pot_get_cheb(self.h, _p(out), C.byref(n))
        return out[: 6 * n.value].reshape(6, n.value)

    def table(self, which):
        n = C.c_int()
        hdr = np.zeros(3)
        lib().orc_pot_get_table(self.h, which, C.byref(n), _p(hdr), None)
        vals = np.zeros(n.value + 3)
        lib().orc_pot_get_table(self.h, which, C.byref(n), _p(hdr), _p(vals))
        return dict(n=n.value, x0=hdr[0], xn=hdr[1], invdx=hdr[2], values=vals)

    def eval(self, which, x):
        """(value, derivative) of phi (0), rho (1) or F (2) at x in the
        current evaluation mode."""
        v, d = C.c_double(), C.c_double()
        lib().orc_pot_eval(self.h, which, float(x), C.byref(v), C.byref(d))
        return v.value, d.value

    def set_cheb3(self, coef):
        """Replace the phi, rho, F value series (3 x n) and rebuild the
        derivative series."""
        c = np.ascontiguousarray(coef, dtype=np.float64).reshape(3, -1)
        lib().orc_pot_set_cheb3(self.h, _p(c), c.shape[1])

    def __del__(self):
        try:
            lib().orc_pot_delete(self.h)
        except Exception:
            pass


class OracleSim:
    """An orc_sim in the reference's slot layout."""

    def __init__(self, handle, pot):
        if not handle:
            raise RuntimeError("oracle simulation could not be created")
        self.h = handle
        self.pot = pot  # keep alive

    @classmethod
    def from_atoms(cls, pot, bounds, xyz, p=None, bf=1.0, periodic=True, itype=0):
        x = np.ascontiguousarray(xyz, dtype=np.float64)
        pp = np.ascontiguousarray(p, dtype=np.float64) if p is not None else None
        h = lib().orc_from_atoms(pot.h, (C.c_double * 3)(*bounds), bf, 1 if periodic else 0, x.shape[0], _p(x),
                                 _p(pp), itype)
        return cls(h, pot)

    @classmethod
    def fcc(cls, pot, n, lat, defgrad=1.0, bf=1.0, periodic=True):
        return cls(lib().orc_create_fcc(pot.h, n, n, n, lat, defgrad, bf, 1 if periodic else 0), pot)

    @classmethod
    def gz(cls, pot, path, bf=1.0, periodic=True):
        return cls(lib().orc_read_gz(pot.h, str(path).encode(), bf, 1 if periodic else 0), pot)

    def info(self):
        nbx = (C.c_int * 3)()
        nb, nt, err = C.c_int(), C.c_int(), C.c_int()
        b, bs = np.zeros(3), np.zeros(3)
        e, st = C.c_double(), C.c_double()
        lib().orc_sim_info(self.h, nbx, C.byref(nb), C.byref(nt), _p(b), _p(bs), C.byref(e), C.byref(st),
                           C.byref(err))
        return dict(nbx=tuple(nbx), nboxes=nb.value, ntot=nt.value, bounds=b, boxsize=bs, e=e.value,
                    stress=st.value, err=err.value)

    def state(self):
        nb = self.info()["nboxes"]
        S = nb * MAXATOMS
        out = dict(natoms=np.zeros(nb, np.int32), id=np.zeros(S, np.int32), itype=np.zeros(S, np.int32),
                   r=np.zeros((S, 3)), p=np.zeros((S, 3)), f=np.zeros((S, 4)), rho=np.zeros(S), fi=np.zeros(S))
        lib().orc_sim_get(self.h, _p(out["natoms"], C.c_int), _p(out["id"], C.c_int), _p(out["itype"], C.c_int),
                          _p(out["r"]), _p(out["p"]), _p(out["f"]), _p(out["rho"]), _p(out["fi"]))
        return out

    def set_rp(self, r, p):
        lib().orc_sim_set_rp(self.h, _p(np.ascontiguousarray(r, np.float64)), _p(np.ascontiguousarray(p, np.float64)))

    def set_state(self, st):
        """Load a slot-layout state dict (natoms, id, itype, r, p)."""
        na = np.ascontiguousarray(st["natoms"], np.int32)
        i = np.ascontiguousarray(st["id"], np.int32)
        t = np.ascontiguousarray(st["itype"], np.int32)
        r = np.ascontiguousarray(st["r"], np.float64)
        p = np.ascontiguousarray(st["p"], np.float64)
        lib().orc_sim_set_state(self.h, _p(na, C.c_int), _p(i, C.c_int), _p(t, C.c_int), _p(r), _p(p))

    def advance_positions(self, dt_internal):
        lib().orc_advance_positions(self.h, dt_internal)

    def advance_velocity(self, dt_internal):
        lib().orc_advance_velocity(self.h, dt_internal)

    def rebox(self):
        lib().orc_rebox_all(self.h)

    def force(self):
        lib().orc_compute_force(self.h)

    def nsteps(self, n, dt=1e-15):
        return lib().orc_n_time_steps(self.h, n, dt)

    def kinetic(self):
        return lib().orc_kinetic_energy(self.h)

    def centrosymmetry(self):
        """Centrosymmetry parameter per slot (A^2; 1000 for atoms with fewer
        than 12 neighbours found)."""
        out = np.zeros(self.info()["nboxes"] * MAXATOMS)
        lib().orc_centrosymmetry(self.h, _p(out))
        return out

    def virial(self):
        """(W, W_pair): the full virial tensor and its pair part (3x3,
        hartree) of the current positions, momenta and last F'."""
        out = np.zeros(18)
        lib().orc_virial_tensor(self.h, _p(out))
        return out[:9].reshape(3, 3), out[9:].reshape(3, 3)

    def __del__(self):
        try:
            lib().orc_sim_free(self.h)
        except Exception:
            pass


# ---------------------------------------------------------------------------
# dumps of oracle/_ref/ref_harness

def read_dump(path):
    """State written by ref_harness: header, natoms, then per occupied slot
    (id, itype, r3, p3, f4, rho, fi) in cell/slot order."""
    with open(path, "rb") as fh:
        raw = fh.read()
    if raw[:8] != b"COMDREF1":
        raise ValueError(f"{path}: not a reference dump")
    nbx0, nbx1, nbx2, nboxes, ntot = struct.unpack_from("<5i", raw, 8)
    dh = struct.unpack_from("<8d", raw, 28)
    off = 28 + 64
    natoms = np.frombuffer(raw, dtype="<i4", count=nboxes, offset=off).copy()
    off += 4 * nboxes
    rec = np.dtype([("id", "<i4"), ("itype", "<i4"), ("v", "<f8", (12,))])
    n = int(natoms.sum())
    recs = np.frombuffer(raw, dtype=rec, count=n, offset=off)
    return dict(nbx=(nbx0, nbx1, nbx2), nboxes=nboxes, ntot=ntot, bounds=np.array(dh[0:3]),
                boxsize=np.array(dh[3:6]), e=dh[6], stress=dh[7], natoms=natoms, id=recs["id"].copy(),
                itype=recs["itype"].copy(), r=recs["v"][:, 0:3].copy(), p=recs["v"][:, 3:6].copy(),
                f=recs["v"][:, 6:10].copy(), rho=recs["v"][:, 10].copy(), fi=recs["v"][:, 11].copy())


def compact(state):
    """Slot-layout state -> occupied slots in cell/slot order (as in a dump)."""
    natoms = state["natoms"]
    nb = natoms.shape[0]
    mask = (np.arange(MAXATOMS)[None, :] < natoms[:, None]).reshape(-1)
    out = dict(natoms=natoms.copy())
    for k in ("id", "itype", "r", "p", "f", "rho", "fi"):
        if k in state:
            out[k] = state[k][: nb * MAXATOMS][mask]
    return out


def ref_binary(cheby=False):
    p = os.path.join(REF_DIR, "ref_cheby" if cheby else "ref_table")
    return p if os.path.exists(p) else None
