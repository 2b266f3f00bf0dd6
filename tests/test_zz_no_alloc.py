"""comd_observe_noalloc (the observables with their reduction scratch in a
static device array, for callers that cannot allocate) against
comd_observe (scratch allocated by the library) through the C-ABI: equal
results, bit for bit, on two calls in a row, the second on a smaller
system, so that the shared scratch is reused with fewer reduction blocks."""
import ctypes as C

import numpy as np
import pytest

import comd_amd as CA
from helpers import seeded_atoms

pytestmark = pytest.mark.gpu

_hiprt = None


def hiprt():
    """The HIP runtime, for the caller-side device buffer of the test."""
    global _hiprt
    if _hiprt is None:
        _hiprt = C.CDLL("libamdhip64.so")
        _hiprt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        _hiprt.hipFree.argtypes = [C.c_void_p]
        _hiprt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _hiprt.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
    return _hiprt


def observe_noalloc(sim, d_out):
    hip, _ = CA.libraries()
    hip.comd_observe_noalloc.argtypes = [C.c_void_p, C.c_void_p]
    rt = hiprt()
    assert rt.hipMemset(d_out, 0xff, C.sizeof(CA.Observables)) == 0  # stale contents must not survive
    sim._check(hip.comd_observe_noalloc(sim.ctx, d_out))
    sim.sync()
    o = CA.Observables()
    assert rt.hipMemcpy(C.addressof(o), d_out, C.sizeof(o), 2) == 0  # device to host
    return o.as_dict()


def same(a, b):
    for k in a:
        if isinstance(a[k], np.ndarray):
            assert np.array_equal(a[k], b[k]), k
        else:
            assert a[k] == b[k], (k, a[k], b[k])


def test_observe_noalloc_equals_observe(gpu, cu500):
    rt = hiprt()
    d_out = C.c_void_p()
    assert rt.hipMalloc(C.byref(d_out), C.sizeof(CA.Observables)) == 0
    try:
        cases = [(cu500, 3.615, 12, "fp64"), (CA.copper_lj(), CA.copper_lj().lat, 6, "mixed")]
        for k, (pot, lat, n, prec) in enumerate(cases):
            print(f"[case] no-alloc observables, call {k}: {n}^3 {prec}", flush=True)
            xyz, p, bounds = seeded_atoms(n, lat, pot.mass, temp=300.0, disp=0.05, seed=40 + k)
            sim = CA.Simulation(pot, bounds, defgrad=0.0, precision=prec)
            sim.bin_atoms(xyz, p)
            sim.rebox()
            ref = sim.observe()
            got = observe_noalloc(sim, d_out)
            print(f"[no-alloc {k}] atoms {got['natoms']} T {got['temperature']:.3f} K "
                  f"W_xx {got['virial'][0, 0]:.6e}", flush=True)
            assert got["natoms"] == len(xyz)
            same(ref, got)
            sim.close()
    finally:
        rt.hipFree(d_out)
