"""comd_amd -- MI355X (gfx950) link-cell force and timestep path of CoMD.

Python view of the C-ABI in ``include/comd_hip.h`` (``libcomd_hip.so``) and
of the setup helpers of ``include/comd_host.h`` (``libcomd_host.so``).  All
force, rebinning and integration work runs in the HIP library; loading fails
with an error when the built libraries are missing (there is no CPU path).

Main objects:

* :class:`EAMPotential` / :class:`LJPotential` -- potentials in the
  reference's form (tables of src-flat/eam.c:130-398, Chebyshev fit of
  src-flat/cheby.c:99-270, LJ parameters of src-flat/ljForce.c:75-86).
* :class:`Simulation` -- one device context: binning, force, rebinning,
  ``nsteps`` (src-flat/timestep.c:114-140), energies, download in the
  reference's slot layout (``c*64+i``).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
POT_DIR = os.path.join(PKG_DIR, "pots")
MAXATOMS = 64

OK = 0
ERRORS = {
    1: "bad argument", 2: "HIP error", 3: "out of device memory", 4: "link cell overflow (64 atoms)",
    5: "atom outside the cell grid", 6: "grid has fewer than one cutoff-wide cell", 7: "no HIP device",
    8: "call out of order",
}
POT_EAM, POT_LJ = 1, 2
EVAL_TABLE, EVAL_CHEB = 0, 1
PREC_FP64, PREC_MIXED = 0, 1
K_REBIN, K_DENSITY, K_EMBED_FORCE, K_REDUCE, K_LIST_BUILD, K_CELL_FALLBACK, K_NCLASS = 0, 1, 2, 3, 4, 5, 6
TUNE_WAVES, TUNE_BLOCKS_PER_CU, TUNE_TABLES_LDS, TUNE_BALANCED, TUNE_GRAPH, TUNE_PAIR_LISTS, TUNE_SKIN = 1, 2, 3, 4, 5, 6, 7
TUNE_LJ_PHYSICAL, TUNE_LIST_GROUPS, TUNE_LIST_ROWS_LDS, TUNE_LIST_THREADS_DIV = 8, 9, 10, 11
TUNE_ADJOINT_LISTS, TUNE_LIST_LANES = 12, 13

# unit constants of src-flat/include/constants.h:61-65
BOHR_PER_ATU_TO_A_PER_S = 0.529 / 2.418e-17
AMU_TO_ME = 1822.83
HARTREE_EV = 27.211386
KB_HARTREE = 3.166811563e-6


class ComdError(RuntimeError):
    def __init__(self, code, msg=""):
        super().__init__(f"comd error {code} ({ERRORS.get(code, '?')}): {msg}")
        self.code = code


_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class Params(C.Structure):
    """struct comd_params of include/comd_hip.h"""
    _fields_ = [
        ("potential", C.c_int), ("eval", C.c_int), ("precision", C.c_int), ("periodic", C.c_int),
        ("cutoff", C.c_double), ("mass", C.c_double), ("lj_sigma", C.c_double), ("lj_epsilon", C.c_double),
        ("tab_n", C.c_int * 3), ("tab_x0", C.c_double * 3), ("tab_xn", C.c_double * 3),
        ("tab_invdx", C.c_double * 3), ("tab_values", _dp * 3),
        ("cheb_n", C.c_int), ("cheb_coef", _dp), ("cheb_x0", C.c_double * 3), ("cheb_xn", C.c_double * 3),
        ("nbx", C.c_int * 3), ("bounds", C.c_double * 3), ("boxsize", C.c_double * 3),
    ]


class PotArray(C.Structure):
    """potentialarray_t (src-flat/include/eamTypes.h:68-74)"""
    _fields_ = [("n", C.c_int), ("x0", C.c_double), ("xn", C.c_double), ("invDx", C.c_double), ("values", _dp)]


class EamPotT(C.Structure):
    """eampotential_t (src-flat/include/eamTypes.h:90-99)"""
    _fields_ = [("cutoff", C.c_double), ("mass", C.c_double), ("force", C.c_void_p), ("destroy", C.c_void_p),
                ("phi", C.POINTER(PotArray)), ("rho", C.POINTER(PotArray)), ("f", C.POINTER(PotArray)),
                ("lat", C.c_double)]


class ChebT(C.Structure):
    """eam_cheby_t (src-flat/include/eamTypes.h:102-112)"""
    _fields_ = [("cutoff", C.c_double), ("mass", C.c_double)] + [
        (n, C.POINTER(PotArray)) for n in ("phi", "rho", "f", "dphi", "drho", "df")] + [("lat", C.c_double)]


class Observables(C.Structure):
    """struct comd_observables of include/comd_hip.h"""
    _fields_ = [("pe", C.c_double), ("ke", C.c_double), ("etot", C.c_double), ("temperature", C.c_double),
                ("virial", C.c_double * 9), ("virial_pair", C.c_double * 9), ("pressure", C.c_double),
                ("stress_ref", C.c_double), ("volume", C.c_double), ("natoms", C.c_int)]

    def as_dict(self):
        out = {k: getattr(self, k) for k in ("pe", "ke", "etot", "temperature", "pressure", "stress_ref", "volume",
                                             "natoms")}
        out["virial"] = np.array(self.virial[:]).reshape(3, 3)
        out["virial_pair"] = np.array(self.virial_pair[:]).reshape(3, 3)
        return out


class SimflatT(C.Structure):
    """simflat_t (src-flat/include/pmdTypes.h:104-141), double precision"""
    _fields_ = [("stateflag", C.c_int), ("nbx", C.c_int * 3), ("nboxes", C.c_int), ("ntot", C.c_int),
                ("bounds", C.c_double * 3), ("boxsize", C.c_double * 3), ("dcenter", _dp), ("natoms", _ip),
                ("id", _ip), ("iType", _ip), ("mass", _dp), ("r", _dp), ("p", _dp), ("f", _dp), ("rho", _dp),
                ("fi", _dp), ("e", C.c_double), ("nAtoms", C.c_int), ("pot", C.c_void_p), ("comment", C.c_char_p),
                ("stress", C.c_double), ("defgrad", C.c_double), ("bf", C.c_double), ("ch_pot", C.POINTER(ChebT))]


class CommandT(C.Structure):
    """command_t (src-flat/include/mycommand.h:60-75)"""
    _fields_ = [("filename", C.c_char * 1024), ("potdir", C.c_char * 1024), ("potname", C.c_char * 1024),
                ("debug", C.c_int), ("doeam", C.c_int), ("periodic", C.c_int), ("usegpu", C.c_int),
                ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("bf", C.c_double), ("lat", C.c_double),
                ("defgrad", C.c_double), ("temp", C.c_double)]


_HIP = None
_HOST = None
_DROPIN = None


def _load():
    global _HIP, _HOST, _DROPIN
    if _HIP is not None:
        return _HIP, _HOST
    hip_path = os.path.join(PKG_DIR, "libcomd_hip.so")
    host_path = os.path.join(PKG_DIR, "libcomd_host.so")
    dropin_path = os.path.join(PKG_DIR, "libcomd_dropin.so")
    for p in (hip_path, dropin_path, host_path):
        if not os.path.exists(p):
            raise ImportError(f"{p} is not built; run make -C comd_amd")
    hip = C.CDLL(hip_path, mode=C.RTLD_GLOBAL)
    host = C.CDLL(host_path, mode=C.RTLD_GLOBAL)
    vp = C.c_void_p
    hip.comd_grid_for.argtypes = [C.POINTER(Params), C.c_double * 3, C.c_double, C.c_double]
    hip.comd_create.argtypes = [C.POINTER(Params), C.POINTER(vp)]
    hip.comd_destroy.argtypes = [vp]
    hip.comd_sizes.argtypes = [vp, _ip, _ip]
    hip.comd_upload.argtypes = [vp, _ip, _ip, _ip, _dp, _dp]
    hip.comd_download.argtypes = [vp, _ip, _ip, _ip, _dp, _dp, _dp, _dp, _dp]
    hip.comd_bin_atoms.argtypes = [vp, C.c_int, _dp, _dp, C.c_int]
    hip.comd_bin_atoms_device.argtypes = [vp, C.c_int, vp, vp, C.c_int]
    hip.comd_rebox.argtypes = [vp]
    hip.comd_force.argtypes = [vp]
    hip.comd_nsteps.argtypes = [vp, C.c_int, C.c_double, _dp]
    hip.comd_steps.argtypes = [vp, C.c_int, C.c_double]
    hip.comd_pair_count.argtypes = [vp, C.POINTER(C.c_longlong)]
    hip.comd_tune.argtypes = [vp, C.c_int, C.c_int]
    hip.comd_download_aux.argtypes = [vp, _ip, _dp]
    hip.comd_energy.argtypes = [vp, _dp, _dp, _dp]
    hip.comd_advance_positions.argtypes = [vp, C.c_double]
    hip.comd_advance_velocity.argtypes = [vp, C.c_double]
    hip.comd_set_precision.argtypes = [vp, C.c_int]
    hip.comd_set_eval.argtypes = [vp, C.c_int]
    hip.comd_sync.argtypes = [vp]
    hip.comd_device_info.argtypes = [_ip, C.c_char_p, C.c_int]
    hip.comd_last_error.argtypes = [vp]
    hip.comd_last_error.restype = C.c_char_p
    hip.comd_timing.argtypes = [vp, C.c_int]
    hip.comd_timing_read.argtypes = [vp, _dp, _ip]
    hip.comd_batch_create.argtypes = [C.c_int, C.POINTER(Params), C.POINTER(vp)]
    hip.comd_batch_info.argtypes = [vp, C.c_int, _ip, _ip, _ip]
    hip.comd_batch_bin_atoms.argtypes = [vp, C.c_int, C.c_int, _dp, _dp, C.c_int]
    hip.comd_batch_upload.argtypes = [vp, C.c_int, _ip, _ip, _ip, _dp, _dp]
    hip.comd_batch_download.argtypes = [vp, C.c_int, _ip, _ip, _ip, _dp, _dp, _dp, _dp, _dp]
    hip.comd_batch_energy.argtypes = [vp, C.c_int, _dp, _dp, _dp]
    hip.comd_batch_step.argtypes = [vp, C.c_int, C.c_double]
    hip.comd_autotune.argtypes = [vp, _ip, _ip, _ip, _dp]
    hip.comd_autotune_lists.argtypes = [vp, _ip, _ip, _ip, _dp]
    hip.comd_list_info.argtypes = [vp, _ip, _ip, _ip, _ip]
    hip.comd_list_occupancy.argtypes = [vp, _ip, _ip, _ip]
    hip.comd_list_lanes.argtypes = [vp, _ip]
    hip.comd_set_cheb.argtypes = [vp, _dp, C.c_int]
    hip.comd_get_cheb.argtypes = [vp, _dp]
    hip.comd_eam_cheb_grad.argtypes = [vp, _dp, C.c_int, _dp, C.c_double, C.c_double, _dp, _dp]
    hip.comd_batch_cheb_grad.argtypes = [vp, _dp, C.c_int, _dp, _dp, C.c_double, _dp, _dp]
    hip.comd_eam_cheb_energy_grad.argtypes = [vp, _dp, C.c_int, _dp, _dp]
    hip.comd_batch_cheb_energy_grad.argtypes = [vp, _dp, C.c_int, _dp, _dp, _dp]
    hip.comd_observe.argtypes = [vp, C.POINTER(Observables)]
    hip.comd_batch_observe.argtypes = [vp, C.c_int, C.POINTER(Observables)]
    hip.comd_run_monitored.argtypes = [vp, C.c_int, C.c_int, C.c_double, _dp]
    hip.comd_centrosymmetry.argtypes = [vp, _dp]
    hip.comd_batch_centrosymmetry.argtypes = [vp, C.c_int, _dp]
    hip.comd_snapshot_save.argtypes = [vp, C.c_char_p]
    hip.comd_snapshot_load.argtypes = [vp, C.c_char_p]
    host.setEamPot.argtypes = [C.c_char_p, C.c_char_p]
    host.setEamPot.restype = C.POINTER(EamPotT)
    host.setChebPot.argtypes = [C.POINTER(EamPotT), C.c_int]
    host.setChebPot.restype = C.POINTER(ChebT)
    host.comd_fcc_positions.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, _dp, C.c_int]
    host.comd_seed_state.argtypes = [C.c_ulonglong, C.c_int, C.c_double, C.c_double, C.c_double, _dp, _dp]
    host.comd_write_gz.argtypes = [C.c_char_p, C.c_int, C.c_double * 3, _dp, _dp]
    host.comd_read_gz.argtypes = [C.c_char_p, _ip, C.c_double * 3, C.POINTER(_dp), C.POINTER(_dp)]
    host.comd_host_error.restype = C.c_char_p
    # the reference's entry points and the drop-in adapters on simflat_t
    sp, pp = C.POINTER(SimflatT), C.c_void_p
    host.create_fcc_lattice.argtypes = [CommandT, pp]
    host.create_fcc_lattice.restype = sp
    host.fromFileGzip.argtypes = [CommandT, pp]
    host.fromFileGzip.restype = sp
    host.getLJPot.restype = pp
    host.destroySimulation.argtypes = [C.POINTER(sp)]
    host.parseCommandLine.argtypes = [C.POINTER(CommandT), C.c_int, C.POINTER(C.c_char_p)]
    host.parseCommandLine.restype = None
    dropin = C.CDLL(dropin_path, mode=C.RTLD_GLOBAL)
    dropin.comd_hip_eam_force.argtypes = [sp]
    dropin.comd_hip_lj_force.argtypes = [sp]
    dropin.comd_hip_nTimeSteps.argtypes = [C.c_int, sp, C.c_double]
    dropin.comd_hip_nTimeSteps.restype = C.c_double
    dropin.comd_hip_reBoxAll.argtypes = [sp]
    dropin.comd_host_error.restype = C.c_char_p
    _HIP, _HOST, _DROPIN = hip, host, dropin
    return hip, host


def libraries():
    """(hip, host) ctypes handles; raises ImportError when not built."""
    return _load()


def dropin():
    """ctypes handle of libcomd_dropin.so (include/comd_dropin.h)."""
    _load()
    return _DROPIN


def device_info():
    """(number of HIP devices, name of the current one)."""
    hip, _ = _load()
    n = C.c_int()
    buf = C.create_string_buffer(256)
    hip.comd_device_info(C.byref(n), buf, 256)
    return n.value, buf.value.decode()


def _ptr(a, ct=C.c_double):
    return a.ctypes.data_as(C.POINTER(ct)) if a is not None else None


# ---------------------------------------------------------------------------
# potentials

class EAMPotential:
    """EAM tables read by the host library (src-flat/eam.c:347-398) and the
    32-term Chebyshev fit of src-flat/cheby.c:99-121."""

    def __init__(self, name="cu500", directory=POT_DIR, ncheb=32):
        _, host = _load()
        p = host.setEamPot(directory.encode(), name.encode())
        if not p:
            raise ComdError(1, f"cannot read EAM potential {directory}/{name}: {host.comd_host_error().decode()}")
        self._p = p
        pot = p.contents
        self.name = name
        self.cutoff = pot.cutoff
        self.mass = pot.mass
        self.lat = pot.lat
        self.tables = []
        for arr in (pot.phi, pot.rho, pot.f):
            a = arr.contents
            vals = np.ctypeslib.as_array(C.cast(C.addressof(a.values.contents) - 8, _dp), shape=(a.n + 3,)).copy()
            self.tables.append(dict(n=a.n, x0=a.x0, xn=a.xn, invdx=a.invDx, values=vals))
        ch = host.setChebPot(p, ncheb).contents
        self.ncheb = ncheb
        self.cheb = np.zeros((6, ncheb))
        self.cheb_dom = np.zeros((3, 2))
        for k, arr in enumerate((ch.phi, ch.rho, ch.f, ch.dphi, ch.drho, ch.df)):
            a = arr.contents
            self.cheb[k] = np.ctypeslib.as_array(a.values, shape=(ncheb,))
            if k < 3:
                self.cheb_dom[k] = (a.x0, a.xn)
        self.kind = POT_EAM

    def fill(self, prm: Params):
        prm.potential = POT_EAM
        prm.cutoff = self.cutoff
        prm.mass = self.mass
        self._keep = [np.ascontiguousarray(t["values"]) for t in self.tables]
        for k, t in enumerate(self.tables):
            prm.tab_n[k] = t["n"]
            prm.tab_x0[k] = t["x0"]
            prm.tab_xn[k] = t["xn"]
            prm.tab_invdx[k] = t["invdx"]
            prm.tab_values[k] = _ptr(self._keep[k])
        self._cheb_keep = np.ascontiguousarray(self.cheb.reshape(-1))
        prm.cheb_n = self.ncheb
        prm.cheb_coef = _ptr(self._cheb_keep)
        for k in range(3):
            prm.cheb_x0[k], prm.cheb_xn[k] = self.cheb_dom[k]

    def with_cheb(self, coef):
        """A copy whose Chebyshev coefficients are replaced by coef (6 x n)."""
        import copy
        q = copy.copy(self)
        q.cheb = np.array(coef, dtype=np.float64).reshape(6, self.ncheb)
        return q


class LJPotential:
    """12-6 Lennard-Jones with settable parameters (src-flat/ljForce.c:75-86
    hard-codes sigma=1.53, eps=0.0085, rc=3 sigma, m=1).  sigma and cutoff
    in A, epsilon in hartree, mass in amu."""

    def __init__(self, sigma=1.53, epsilon=0.0085, cutoff=None, mass=1.0):
        self.sigma = float(sigma)
        self.epsilon = float(epsilon)
        self.cutoff = float(cutoff) if cutoff is not None else 3.0 * self.sigma
        self.mass = float(mass)
        self.kind = POT_LJ

    def fill(self, prm: Params):
        prm.potential = POT_LJ
        prm.cutoff = self.cutoff
        prm.mass = self.mass
        prm.lj_sigma = self.sigma
        prm.lj_epsilon = self.epsilon


def copper_lj():
    """Lennard-Jones copper of the 32 000-atom benchmark case: sigma 2.315 A,
    eps 0.167 eV, rc 2.5 sigma, Cu mass; lattice sqrt(2) 2^(1/6) sigma."""
    s = 2.315
    pot = LJPotential(sigma=s, epsilon=0.167 / HARTREE_EV, cutoff=2.5 * s, mass=63.55)
    pot.lat = np.sqrt(2.0) * 2.0 ** (1.0 / 6.0) * s
    return pot


# ---------------------------------------------------------------------------
# initial states

def fcc_positions(nx, ny, nz, lat, defgrad=1.0):
    """Atom positions in the order of create_fcc_lattice (src-flat/ic_fcc.c:115-132)."""
    _, host = _load()
    n = host.comd_fcc_positions(nx, ny, nz, lat, defgrad, None, 0)
    xyz = np.zeros((n, 3))
    host.comd_fcc_positions(nx, ny, nz, lat, defgrad, _ptr(xyz), n)
    return xyz


def seed_state(xyz, mass_amu, temp, disp, seed):
    """Seeded Gaussian displacements and Maxwell-Boltzmann momenta (zero net
    momentum); returns (xyz_displaced, p)."""
    _, host = _load()
    x = np.ascontiguousarray(xyz, dtype=np.float64).copy()
    p = np.zeros_like(x)
    host.comd_seed_state(int(seed), x.shape[0], float(mass_amu), float(temp), float(disp), _ptr(x), _ptr(p))
    return x, p


def write_gz(path, bounds, xyz, p=None):
    _, host = _load()
    b = (C.c_double * 3)(*bounds)
    x = np.ascontiguousarray(xyz, dtype=np.float64)
    pp = np.ascontiguousarray(p, dtype=np.float64) if p is not None else None
    rc = host.comd_write_gz(str(path).encode(), x.shape[0], b, _ptr(x), _ptr(pp))
    if rc:
        raise ComdError(rc, host.comd_host_error().decode())


def read_gz(path):
    """Reference gz input (src-flat/read.c:155-241) -> (bounds, xyz, p)."""
    _, host = _load()
    n = C.c_int()
    b = (C.c_double * 3)()
    xp, pp = _dp(), _dp()
    rc = host.comd_read_gz(str(path).encode(), C.byref(n), b, C.byref(xp), C.byref(pp))
    if rc:
        raise ComdError(rc, host.comd_host_error().decode())
    m = n.value
    xyz = np.ctypeslib.as_array(xp, shape=(max(m, 1), 3))[:m].copy()
    p = np.ctypeslib.as_array(pp, shape=(max(m, 1), 3))[:m].copy()
    libc = C.CDLL(None)
    libc.free(C.cast(xp, C.c_void_p))
    libc.free(C.cast(pp, C.c_void_p))
    return list(b), xyz, p


# ---------------------------------------------------------------------------

class Simulation:
    """One device context (comd_create) for a potential on a link-cell grid.

    The grid follows allocDomains (src-flat/domains.c:155-200): pass
    ``defgrad`` as the reference would see it (the FCC path keeps the
    stretch, the file path uses 0)."""

    def __init__(self, pot, bounds, defgrad=1.0, bf=1.0, periodic=True, precision="fp64", evaluation="table"):
        hip, _ = _load()
        self._hip = hip
        self.pot = pot
        prm = Params()
        pot.fill(prm)
        prm.periodic = 1 if periodic else 0
        prm.precision = PREC_MIXED if precision == "mixed" else PREC_FP64
        prm.eval = EVAL_CHEB if evaluation == "cheby" else EVAL_TABLE
        rc = hip.comd_grid_for(C.byref(prm), (C.c_double * 3)(*bounds), float(defgrad), float(bf))
        if rc:
            raise ComdError(rc, "grid")
        self.prm = prm
        self.nbx = tuple(prm.nbx)
        self.ncells = int(np.prod(self.nbx))
        self.bounds = tuple(prm.bounds)
        self.boxsize = tuple(prm.boxsize)
        ctx = C.c_void_p()
        rc = hip.comd_create(C.byref(prm), C.byref(ctx))
        if rc:
            raise ComdError(rc, hip.comd_last_error(None).decode())
        self.ctx = ctx

    def close(self):
        if getattr(self, "ctx", None):
            self._hip.comd_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, rc):
        if rc:
            raise ComdError(rc, self._hip.comd_last_error(self.ctx).decode())

    @property
    def nslots(self):
        return self.ncells * MAXATOMS

    def bin_atoms(self, xyz, p=None, itype=0):
        x = np.ascontiguousarray(xyz, dtype=np.float64)
        pp = np.ascontiguousarray(p, dtype=np.float64) if p is not None else None
        self._check(self._hip.comd_bin_atoms(self.ctx, x.shape[0], _ptr(x), _ptr(pp), int(itype)))

    @staticmethod
    def _slot_arrays(ncells, natoms, ids, itype, r, p):
        """Contiguous copies of a slot-layout state, with the shapes the C
        side reads (it reads ncells counts, ncells*64 ids and types and
        ncells*64 x 3 positions and momenta)."""
        S = ncells * MAXATOMS
        na = np.ascontiguousarray(natoms, dtype=np.int32)
        i = np.ascontiguousarray(ids, dtype=np.int32)
        t = np.ascontiguousarray(itype, dtype=np.int32)
        rr = np.ascontiguousarray(r, dtype=np.float64)
        pp = np.ascontiguousarray(p, dtype=np.float64)
        for name, a, shape in (("natoms", na, (ncells,)), ("ids", i, (S,)), ("itype", t, (S,)), ("r", rr, (S, 3)),
                               ("p", pp, (S, 3))):
            if a.shape != shape:
                raise ValueError(f"{name} has shape {a.shape}, expected {shape}")
        return na, i, t, rr, pp

    def upload(self, natoms, ids, itype, r, p):
        na, i, t, rr, pp = self._slot_arrays(self.ncells, natoms, ids, itype, r, p)
        self._check(self._hip.comd_upload(self.ctx, _ptr(na, C.c_int), _ptr(i, C.c_int), _ptr(t, C.c_int),
                                          _ptr(rr), _ptr(pp)))

    def download(self):
        """Slot-layout state: natoms (C), id, itype (S), r, p (S,3), f (S,4), rho, fi (S)."""
        S, Cn = self.nslots, self.ncells
        out = dict(natoms=np.zeros(Cn, np.int32), id=np.zeros(S, np.int32), itype=np.zeros(S, np.int32),
                   r=np.zeros((S, 3)), p=np.zeros((S, 3)), f=np.zeros((S, 4)), rho=np.zeros(S), fi=np.zeros(S))
        self._check(self._hip.comd_download(
            self.ctx, _ptr(out["natoms"], C.c_int), _ptr(out["id"], C.c_int), _ptr(out["itype"], C.c_int),
            _ptr(out["r"]), _ptr(out["p"]), _ptr(out["f"]), _ptr(out["rho"]), _ptr(out["fi"])))
        return out

    def download_aux(self):
        """Per-slot neighbour counts of the last density/LJ pass and per-atom virial."""
        nnb = np.zeros(self.nslots, np.int32)
        vir = np.zeros(self.nslots)
        self._check(self._hip.comd_download_aux(self.ctx, _ptr(nnb, C.c_int), _ptr(vir)))
        return nnb, vir

    def rebox(self):
        self._check(self._hip.comd_rebox(self.ctx))

    def force(self):
        self._check(self._hip.comd_force(self.ctx))

    def nsteps(self, n, dt=1.0e-15):
        pe = C.c_double()
        self._check(self._hip.comd_nsteps(self.ctx, int(n), float(dt), C.byref(pe)))
        return pe.value

    def steps(self, n, dt=1.0e-15):
        """n DKD steps without the trailing force of nTimeSteps; asynchronous."""
        self._check(self._hip.comd_steps(self.ctx, int(n), float(dt)))

    def advance_positions(self, dt_internal):
        self._check(self._hip.comd_advance_positions(self.ctx, float(dt_internal)))

    def advance_velocity(self, dt_internal):
        self._check(self._hip.comd_advance_velocity(self.ctx, float(dt_internal)))

    def energy(self):
        """(potential energy, kinetic energy, virial 'stress') in hartree."""
        pe, ke, st = C.c_double(), C.c_double(), C.c_double()
        self._check(self._hip.comd_energy(self.ctx, C.byref(pe), C.byref(ke), C.byref(st)))
        return pe.value, ke.value, st.value

    def tune(self, waves_per_block=None, blocks_per_cu=None, tables_in_lds=None, balanced=None, graph=None,
             pair_lists=None, skin=None, lj_physical=None, list_groups=None, list_rows_lds=None,
             list_threads_div=None, adjoint_lists=None, list_lanes=None):
        """Launch shape of the force kernels, step replay as a HIP graph,
        pair-list replay in the EAM embedding-force pass, the Verlet skin in A
        (0 turns the neighbour lists off) and the Lennard-Jones force sign
        (lj_physical: atoms move down the potential), the list passes' shape
        (list_groups, list_rows_lds, list_threads_div, list_lanes: lanes per
        home atom, 0 by size) and whether the adjoint's pair passes use the
        lists (adjoint_lists) (comd_tune)."""
        for key, val in ((TUNE_WAVES, waves_per_block), (TUNE_BLOCKS_PER_CU, blocks_per_cu),
                         (TUNE_TABLES_LDS, tables_in_lds), (TUNE_BALANCED, balanced), (TUNE_GRAPH, graph),
                         (TUNE_PAIR_LISTS, pair_lists), (TUNE_SKIN, None if skin is None else round(1000 * skin)),
                         (TUNE_LJ_PHYSICAL, lj_physical), (TUNE_LIST_GROUPS, list_groups),
                         (TUNE_LIST_ROWS_LDS, list_rows_lds), (TUNE_LIST_THREADS_DIV, list_threads_div),
                         (TUNE_ADJOINT_LISTS, adjoint_lists), (TUNE_LIST_LANES, list_lanes)):
            if val is not None:
                self._check(self._hip.comd_tune(self.ctx, key, int(val)))

    def list_info(self):
        """Neighbour lists: dict(active, valid, builds, kmax) (comd_list_info)."""
        a, v, b, k = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        self._check(self._hip.comd_list_info(self.ctx, C.byref(a), C.byref(v), C.byref(b), C.byref(k)))
        return dict(active=a.value, valid=v.value, builds=b.value, kmax=k.value)

    def list_occupancy(self):
        """Neighbour-list passes: dict(blocks_per_cu, lds_bytes, threads)
        (comd_list_occupancy); lists of two (density or LJ, embedding force)."""
        b, l, t = (C.c_int * 2)(), (C.c_int * 2)(), C.c_int()
        self._check(self._hip.comd_list_occupancy(self.ctx, b, l, C.byref(t)))
        return dict(blocks_per_cu=list(b), lds_bytes=list(l), threads=t.value)

    def list_lanes(self):
        """Lanes per home atom of the list passes (comd_list_lanes); 0 without lists."""
        n = C.c_int()
        self._check(self._hip.comd_list_lanes(self.ctx, C.byref(n)))
        return n.value

    def autotune_lists(self):
        """Time the list passes over group shapes, table rows in LDS or not
        and home atoms per thread, and keep the fastest (comd_autotune_lists);
        returns dict(groups, rows_lds, threads_div, ms)."""
        g, r, d = C.c_int(), C.c_int(), C.c_int()
        ms = C.c_double()
        self._check(self._hip.comd_autotune_lists(self.ctx, C.byref(g), C.byref(r), C.byref(d), C.byref(ms)))
        return dict(groups=g.value, rows_lds=r.value, threads_div=d.value, ms=ms.value)

    def autotune(self):
        """Time the force passes that run over launch shapes and keep the
        fastest (comd_autotune: the list passes' shape when the context has
        lists); returns (waves, blocks_per_cu, balanced, ms)."""
        w, b, d = C.c_int(), C.c_int(), C.c_int()
        ms = C.c_double()
        self._check(self._hip.comd_autotune(self.ctx, C.byref(w), C.byref(b), C.byref(d), C.byref(ms)))
        return w.value, b.value, d.value, ms.value

    def pair_count(self):
        """Half-pairs inside the cutoff found by the last density/LJ pass."""
        n = C.c_longlong()
        self._check(self._hip.comd_pair_count(self.ctx, C.byref(n)))
        return n.value

    def set_precision(self, precision):
        self._check(self._hip.comd_set_precision(self.ctx, PREC_MIXED if precision == "mixed" else PREC_FP64))

    def set_eval(self, evaluation):
        self._check(self._hip.comd_set_eval(self.ctx, EVAL_CHEB if evaluation == "cheby" else EVAL_TABLE))

    def sync(self):
        self._check(self._hip.comd_sync(self.ctx))

    def timing(self, on=True):
        self._check(self._hip.comd_timing(self.ctx, 1 if on else 0))

    def timing_read(self):
        ms = (C.c_double * K_NCLASS)()
        ln = (C.c_int * K_NCLASS)()
        self._check(self._hip.comd_timing_read(self.ctx, ms, ln))
        return list(ms), list(ln)

    # ---- Chebyshev coefficients and their adjoint

    def observe(self):
        """Force at the current positions, then PE, KE, temperature, the full
        virial tensor (pair + embedding + kinetic) and its pair part,
        pressure and the reference's scalar stress (comd_observe)."""
        o = Observables()
        self._check(self._hip.comd_observe(self.ctx, C.byref(o)))
        return o.as_dict()

    def run_monitored(self, nblocks, block_steps, dt=1.0e-15):
        """PE and KE after a force at the start and after each of nblocks
        nTimeSteps(block_steps) calls, recorded on the device
        (comd_run_monitored); returns two arrays of nblocks + 1."""
        out = np.zeros(2 * (nblocks + 1))
        self._check(self._hip.comd_run_monitored(self.ctx, int(nblocks), int(block_steps), float(dt), _ptr(out)))
        return out[0::2].copy(), out[1::2].copy()

    def centrosymmetry(self):
        """Centrosymmetry parameter per slot (A^2; 1000 for atoms with fewer
        than 12 neighbours in reach), src-viz's defect measure
        (comd_centrosymmetry)."""
        out = np.zeros(self.nslots)
        self._check(self._hip.comd_centrosymmetry(self.ctx, _ptr(out)))
        return out

    def snapshot_save(self, path):
        """The whole device state, lists included, to a file (comd_snapshot_save)."""
        self._check(self._hip.comd_snapshot_save(self.ctx, str(path).encode()))

    def snapshot_load(self, path):
        """Restore a snapshot into this context (same parameters); the run
        then continues bit for bit (comd_snapshot_load)."""
        self._check(self._hip.comd_snapshot_load(self.ctx, str(path).encode()))

    def set_cheb(self, coef):
        """Replace the phi, rho, F value series (3 x n) and rebuild the
        derivative series (comd_set_cheb)."""
        c = np.ascontiguousarray(coef, dtype=np.float64).reshape(3, -1)
        self._check(self._hip.comd_set_cheb(self.ctx, _ptr(c), c.shape[1]))

    def get_cheb(self):
        """The six series in use: phi, rho, F, dphi, drho, dF (6 x n)."""
        n = self.prm.cheb_n
        out = np.zeros((6, n))
        self._check(self._hip.comd_get_cheb(self.ctx, _ptr(out)))
        return out

    def cheb_grad(self, coef, f_target, e_target, w_e=1.0):
        """Force-matching loss sum_i |f_i - t_i|^2 / (3N) + w_e (E - E_t)^2 at
        the value coefficients `coef` (3 x n) and its gradient (3 x n), from
        the adjoint kernels (comd_eam_cheb_grad).  f_target: slot layout
        (S, 3) or (S, 4) as download()['f'], or None to reuse the target of
        the previous call (kept in device memory)."""
        c = np.ascontiguousarray(coef, dtype=np.float64).reshape(3, -1)
        ft = None
        if f_target is not None:
            ft = np.ascontiguousarray(np.asarray(f_target, dtype=np.float64)[:, :3])
            if ft.shape[0] != self.nslots:
                raise ValueError(f"f_target has {ft.shape[0]} rows, the context {self.nslots} slots")
        loss = C.c_double()
        grad = np.zeros_like(c)
        self._check(self._hip.comd_eam_cheb_grad(self.ctx, _ptr(c), c.shape[1], _ptr(ft), float(e_target),
                                                 float(w_e), C.byref(loss), _ptr(grad)))
        return loss.value, grad

    def cheb_energy_grad(self, coef):
        """Potential energy E and its gradient dE/dc (3 x n) at the value
        coefficients `coef` (comd_eam_cheb_energy_grad)."""
        c = np.ascontiguousarray(coef, dtype=np.float64).reshape(3, -1)
        e = C.c_double()
        grad = np.zeros_like(c)
        self._check(self._hip.comd_eam_cheb_energy_grad(self.ctx, _ptr(c), c.shape[1], C.byref(e), _ptr(grad)))
        return e.value, grad


class Batch(Simulation):
    """Independent replicas in one context (comd_batch_create): every kernel
    of a step covers all of them in one launch, and steps replay a captured
    HIP graph.  `replicas` is a list of dicts with keys pot, bounds and
    optionally defgrad (grid rule of allocDomains), evaluation, periodic."""

    def __init__(self, replicas, precision="mixed", bf=1.0):
        hip, _ = _load()
        self._hip = hip
        n = len(replicas)
        arr = (Params * n)()
        self.pots = []
        self.nbx_list = []
        for r, rep in enumerate(replicas):
            prm = arr[r]
            rep["pot"].fill(prm)
            prm.periodic = 1 if rep.get("periodic", True) else 0
            prm.precision = PREC_MIXED if precision == "mixed" else PREC_FP64
            prm.eval = EVAL_CHEB if rep.get("evaluation", "table") == "cheby" else EVAL_TABLE
            rc = hip.comd_grid_for(C.byref(prm), (C.c_double * 3)(*rep["bounds"]), float(rep.get("defgrad", 0.0)),
                                   float(bf))
            if rc:
                raise ComdError(rc, f"grid of replica {r}")
            self.pots.append(rep["pot"])
            self.nbx_list.append(tuple(prm.nbx))
        self._params = arr
        ctx = C.c_void_p()
        rc = hip.comd_batch_create(n, arr, C.byref(ctx))
        if rc:
            raise ComdError(rc, hip.comd_last_error(None).decode())
        self.ctx = ctx
        self.nrep = n
        self.prm = arr[0]
        cells = C.c_int()
        hip.comd_sizes(ctx, C.byref(cells), None)
        self.ncells = cells.value

    def info(self, r):
        """(first global cell, cells, atoms) of replica r."""
        b, nc, na = C.c_int(), C.c_int(), C.c_int()
        self._check(self._hip.comd_batch_info(self.ctx, r, C.byref(b), C.byref(nc), C.byref(na)))
        return b.value, nc.value, na.value

    def bin_atoms(self, r, xyz, p=None, itype=0):
        x = np.ascontiguousarray(xyz, dtype=np.float64)
        pp = np.ascontiguousarray(p, dtype=np.float64) if p is not None else None
        self._check(self._hip.comd_batch_bin_atoms(self.ctx, r, x.shape[0], _ptr(x), _ptr(pp), int(itype)))

    def upload(self, r, natoms, ids, itype, rr, p):
        """Slot-layout state of replica r (comd_batch_upload)."""
        _, Cn, _ = self.info(r)
        na, i, t, x, pp = self._slot_arrays(Cn, natoms, ids, itype, rr, p)
        self._check(self._hip.comd_batch_upload(self.ctx, r, _ptr(na, C.c_int), _ptr(i, C.c_int), _ptr(t, C.c_int),
                                                _ptr(x), _ptr(pp)))

    def download(self, r):
        """Slot-layout state of replica r (as Simulation.download)."""
        _, Cn, _ = self.info(r)
        S = Cn * MAXATOMS
        out = dict(natoms=np.zeros(Cn, np.int32), id=np.zeros(S, np.int32), itype=np.zeros(S, np.int32),
                   r=np.zeros((S, 3)), p=np.zeros((S, 3)), f=np.zeros((S, 4)), rho=np.zeros(S), fi=np.zeros(S))
        self._check(self._hip.comd_batch_download(
            self.ctx, r, _ptr(out["natoms"], C.c_int), _ptr(out["id"], C.c_int), _ptr(out["itype"], C.c_int),
            _ptr(out["r"]), _ptr(out["p"]), _ptr(out["f"]), _ptr(out["rho"]), _ptr(out["fi"])))
        return out

    def observe_of(self, r):
        """comd_batch_observe of replica r (as Simulation.observe)."""
        o = Observables()
        self._check(self._hip.comd_batch_observe(self.ctx, r, C.byref(o)))
        return o.as_dict()

    def centrosymmetry_of(self, r):
        """comd_batch_centrosymmetry of replica r (as Simulation.centrosymmetry)."""
        _, Cn, _ = self.info(r)
        out = np.zeros(Cn * MAXATOMS)
        self._check(self._hip.comd_batch_centrosymmetry(self.ctx, r, _ptr(out)))
        return out

    def energy_of(self, r):
        pe, ke, st = C.c_double(), C.c_double(), C.c_double()
        self._check(self._hip.comd_batch_energy(self.ctx, r, C.byref(pe), C.byref(ke), C.byref(st)))
        return pe.value, ke.value, st.value

    def cheb_grad(self, coef, f_targets, e_targets, w_e=1.0):
        """Loss sum_r [sum_{i in r} |f_i - t_i|^2 / (3 N_r) + w_e (E_r -
        E_t,r)^2] over the replicas and its gradient (3 x n) at the value
        coefficients `coef`, one launch per adjoint stage
        (comd_batch_cheb_grad).  f_targets: per replica its slot-layout
        target (as download(r)['f']), or None to reuse the previous one;
        e_targets: per replica."""
        c = np.ascontiguousarray(coef, dtype=np.float64).reshape(3, -1)
        ft = None
        if f_targets is not None:
            ft = np.zeros((self.ncells * MAXATOMS, 3))
            for r, t in enumerate(f_targets):
                b, nc, _ = self.info(r)
                ft[b * MAXATOMS:(b + nc) * MAXATOMS] = np.asarray(t, dtype=np.float64)[:, :3]
        et = np.ascontiguousarray(e_targets, dtype=np.float64)
        if et.shape != (self.nrep,):
            raise ValueError(f"e_targets needs {self.nrep} values")
        loss = C.c_double()
        grad = np.zeros_like(c)
        self._check(self._hip.comd_batch_cheb_grad(self.ctx, _ptr(c), c.shape[1], _ptr(ft), _ptr(et), float(w_e),
                                                   C.byref(loss), _ptr(grad)))
        return loss.value, grad

    def cheb_energy_grad(self, coef, weights=None):
        """Energies E_r of the replicas and the gradient of sum_r w_r E_r
        (weights None: all 1) with respect to the value coefficients `coef`
        (comd_batch_cheb_energy_grad)."""
        c = np.ascontiguousarray(coef, dtype=np.float64).reshape(3, -1)
        w = None if weights is None else np.ascontiguousarray(weights, dtype=np.float64)
        if w is not None and w.shape != (self.nrep,):
            raise ValueError(f"weights needs {self.nrep} values")
        e = np.zeros(self.nrep)
        grad = np.zeros_like(c)
        self._check(self._hip.comd_batch_cheb_energy_grad(self.ctx, _ptr(c), c.shape[1], _ptr(w), _ptr(e),
                                                          _ptr(grad)))
        return e, grad


def fcc_simulation(pot, n, lat=None, temp=0.0, disp=0.0, seed=1, precision="fp64", evaluation="table",
                   defgrad=1.0, bf=1.0, periodic=True):
    """FCC n^3 crystal binned on the GPU in create_fcc_lattice order, with
    optional seeded displacements and momenta.  Returns (sim, xyz, p)."""
    lat = lat if lat is not None else pot.lat
    xyz = fcc_positions(n, n, n, lat, defgrad)
    p = np.zeros_like(xyz)
    if temp > 0 or disp > 0:
        xyz, p = seed_state(xyz, pot.mass, temp, disp, seed)
    bounds = (n * lat * defgrad, n * lat, n * lat)
    sim = Simulation(pot, bounds, defgrad=defgrad, bf=bf, periodic=periodic, precision=precision,
                     evaluation=evaluation)
    sim.bin_atoms(xyz, p, itype=1)
    return sim, xyz, p
