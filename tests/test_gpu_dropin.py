"""The drop-in boundary as the reference's callers use it.

* The reference's own program (its src-flat sources compiled with
  oracle/ref_harness.c into oracle/_ref/ref_dropin, linked with
  libcomd_dropin.so) with comd_hip_eam_force / comd_hip_lj_force stored in
  the potential's force slot, so that the reference's computeForce and
  nTimeSteps (src-flat/timestep.c:110-140) call the GPU force; and with
  comd_hip_nTimeSteps in place of nTimeSteps.  Its dumps are compared with
  the fixtures the unmodified reference wrote.
* The same entry points through libcomd_host on a simflat_t built by
  setEamPot / fromFileGzip / setChebPot: the host arrays after each call
  against the fixtures.
* The command-line driver (comd_amd/comd, do_compute_work of
  src-flat/docompute.c:166-218) against the energies the reference prints
  for its data/2.inp.gz and data/8k.inp.gz inputs.
"""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

import comd_amd as CA
import oracle_py as O
from helpers import GOLDEN, POTS, ROOT, force_rel_error, scalar_rel_error

pytestmark = pytest.mark.gpu

DROPIN = os.path.join(ROOT, "oracle", "_ref", "ref_dropin")
COMD = os.path.join(ROOT, "comd_amd", "comd")
LJ_ARGS = ["--lj", repr(2.315), repr(0.167 / 27.211386), repr(2.5 * 2.315), repr(63.55)]


def compare_dump(d, fx, m, what, rtol_r=1e-10):
    """A dump of the reference program against fixture state m."""
    assert np.array_equal(d["natoms"], fx[f"s{m}_natoms"]), f"{what}: cell counts"
    assert np.array_equal(d["id"], fx[f"s{m}_id"]), f"{what}: slot order"
    ef = force_rel_error(d["f"], fx[f"s{m}_f"]).max()
    ee = scalar_rel_error(d["f"][:, 3], fx[f"s{m}_f"][:, 3]).max()
    dr = np.abs(d["r"] - fx[f"s{m}_r"]).max()
    dp = np.abs(d["p"] - fx[f"s{m}_p"]).max() / max(np.abs(fx[f"s{m}_p"]).max(), 1e-300)
    de = abs(d["e"] - float(fx[f"s{m}_e"])) / abs(float(fx[f"s{m}_e"]))
    print(f"[{what} state {m}] force {ef:.2e} energy {ee:.2e} r {dr:.2e} p {dp:.2e} E {de:.2e}", flush=True)
    assert ef <= 1e-10 and ee <= 1e-10 and dr <= rtol_r and dp <= 1e-10 and de <= 1e-12


@pytest.mark.skipif(not os.path.exists(DROPIN), reason="oracle/_ref/ref_dropin is built only next to the reference tree")
@pytest.mark.parametrize("mode", ["--hip-force", "--hip-steps"])
@pytest.mark.parametrize("case", ["cu6_table", "lj6"])
def test_reference_program_with_gpu_force_slot(gpu, tmp_path, mode, case):
    fx = np.load(os.path.join(GOLDEN, f"{case}.npz"))
    inp = os.path.join(GOLDEN, "data", "cu6.inp.gz" if case == "cu6_table" else "lj6.inp.gz")
    pot = ["--eam", POTS, "cu500"] if case == "cu6_table" else LJ_ARGS
    steps, dumps = int(fx["steps"]), int(fx["dumps"])
    prefix = str(tmp_path / case)
    cmd = [DROPIN] + pot + ["--input", inp, "--steps", str(steps), "--dumps", str(dumps), mode, "--out", prefix]
    print("[case] " + " ".join(os.path.relpath(c, ROOT) if c.startswith(ROOT) else c for c in cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    for m in range(dumps + 1):
        compare_dump(O.read_dump(f"{prefix}.{m}.bin"), fx, m, f"{case} {mode}")


def host_arrays(s):
    nb = s.nboxes
    S = nb * 64
    out = dict(natoms=np.ctypeslib.as_array(s.natoms, shape=(nb,)).copy(),
               id=np.ctypeslib.as_array(s.id, shape=(S,)).copy(),
               r=np.ctypeslib.as_array(s.r, shape=(S, 3)).copy(),
               p=np.ctypeslib.as_array(s.p, shape=(S, 3)).copy(),
               f=np.ctypeslib.as_array(s.f, shape=(S, 4)).copy(),
               rho=np.ctypeslib.as_array(s.rho, shape=(S,)).copy(),
               fi=np.ctypeslib.as_array(s.fi, shape=(S,)).copy())
    m = (np.arange(64)[None, :] < out["natoms"][:, None]).reshape(-1)
    comp = {k: (v[m] if k != "natoms" else v) for k, v in out.items()}
    comp["e"] = s.e
    comp["stress"] = s.stress
    return comp


def test_adapters_on_a_reference_simflat(gpu):
    """setEamPot -> fromFileGzip (command_t by value) -> setChebPot, then
    comd_hip_eam_force, comd_hip_nTimeSteps and comd_hip_reBoxAll on the
    simflat_t: the host arrays after each call equal the reference's."""
    _, host = CA.libraries()
    d = CA.dropin()
    fx = np.load(os.path.join(GOLDEN, "cu6_table.npz"))
    pot = host.setEamPot(POTS.encode(), b"cu500")
    cmd = CA.CommandT()
    cmd.filename = os.path.join(GOLDEN, "data", "cu6.inp.gz").encode()
    cmd.bf, cmd.defgrad, cmd.periodic = 1.0, 1.0, 1
    sp = host.fromFileGzip(cmd, C.cast(pot, C.c_void_p))
    assert sp
    s = sp.contents
    s.ch_pot = host.setChebPot(pot, 32)
    assert tuple(s.nbx) == tuple(fx["nbx"])
    assert d.comd_hip_reBoxAll(sp) == 0
    assert d.comd_hip_eam_force(sp) == 0
    compare_dump(host_arrays(s), fx, 0, "adapter force")
    h = host_arrays(s)
    np.testing.assert_allclose(h["rho"], fx["s0_rho"], rtol=1e-10)
    np.testing.assert_allclose(h["fi"], fx["s0_fi"], rtol=1e-10)
    assert abs(h["stress"] - float(fx["s0_stress"])) <= 1e-10 * abs(float(fx["s0_stress"]))
    steps = int(fx["steps"])
    for m in range(1, int(fx["dumps"]) + 1):
        e = d.comd_hip_nTimeSteps(steps, sp, 1.0e-15)
        assert d.comd_host_status() == 0, d.comd_host_error()
        compare_dump(host_arrays(s), fx, m, "adapter nTimeSteps")
        assert e == s.e
    host.destroySimulation(C.byref(sp))


def run_driver(args):
    r = subprocess.run([COMD] + args, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    return [float(m.group(2)) for m in re.finditer(r"^\s*(\d+)\s+(-?[0-9.eE+-]+) computed", r.stdout, flags=re.M)]


# The reference's printed s->e at steps 0, 10 and 20 (gcc, -O3 -DDOUBLE,
# -DGZIPSUPPORT [, -DUSE_CHEBY], x86-64; data/2.inp.gz, data/8k.inp.gz)
REFERENCE_TRACES = {
    ("cu500", "2.inp.gz", "table"): (-0.069800808279795789613, -0.072492997606632220142, -0.075811819287074477813),
    ("ag", "8k.inp.gz", "table"): (-852.31072722864178104, -852.67331452789107971, -853.64647892137020335),
    ("ag", "8k.inp.gz", "cheby"): (-852.27037691173302392, -852.63264506536700083, -853.60499685348986532),
}


@pytest.mark.parametrize("pot,inp,ev", sorted(REFERENCE_TRACES))
def test_driver_energy_trace(gpu, pot, inp, ev):
    """comd_amd/comd -e -p POT -f INPUT: the energy printed after the initial
    force and after each nTimeSteps(10) call, against the reference's."""
    args = ["-e", "-d", POTS, "-p", pot, "-f", os.path.join(GOLDEN, "data", inp), "--niter", "2", "--nsteps", "10",
            "--eval", ev]
    e = run_driver(args)
    ref = REFERENCE_TRACES[(pot, inp, ev)]
    print(f"[driver {pot} {inp} {ev}] {e} reference {ref}", flush=True)
    assert len(e) == 3
    for a, b in zip(e, ref):
        assert abs(a - b) <= 1e-11 * abs(b)
