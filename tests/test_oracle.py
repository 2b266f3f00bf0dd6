"""CPU checks that pin the oracle (oracle/liboracle.so).

Two kinds of evidence, so that a plausible mistake in the restatement (a
dropped term, a wrong sign or index, a swapped predicate) fails at least one:

* golden fixtures written by the compiled reference itself
  (tests/golden/make_golden.py): per-slot states after the initial force and
  after further nTimeSteps calls, table headers, guard slots and Chebyshev
  coefficients -- compared bit for bit in table mode;
* what the mathematics fixes independently of any code path: the perfect-FCC
  lattice sum, the two-atom EAM and Lennard-Jones closed forms, the
  interpolation property of the Chebyshev fit at its nodes, the derivative
  series against finite differences, Clenshaw against numpy's Chebyshev
  series, Newton's third law, translation invariance, free flight and the
  time reversibility of velocity Verlet, and floor semantics at cell faces.
"""
import math
import os

import numpy as np
import pytest

import comd_amd as CA
import oracle_py as O
from helpers import GOLDEN, POTS, DT_INTERNAL

LJ_SIGMA, LJ_EPS, LJ_CUT, LJ_MASS = 2.315, 0.167 / 27.211386, 2.5 * 2.315, 63.55
HARTREE_EV = 27.211386


def golden(name):
    return np.load(os.path.join(GOLDEN, name))


def data(name):
    return os.path.join(GOLDEN, "data", name)


@pytest.fixture(scope="module")
def cu():
    return O.OraclePot.eam(POTS, "cu500")


def absolute_positions(sim):
    """id -> absolute position (relative position + lower cell corner,
    src-flat/domains.c:191-197,280)."""
    info = sim.info()
    st = sim.state()
    nbx = info["nbx"]
    natoms = st["natoms"]
    out = np.zeros((info["ntot"], 3))
    for c in np.flatnonzero(natoms):
        ix, iy, iz = c % nbx[0], (c // nbx[0]) % nbx[1], c // (nbx[0] * nbx[1])
        corner = info["boxsize"] * np.array([ix, iy, iz])
        for i in range(natoms[c]):
            o = c * O.MAXATOMS + i
            out[st["id"][o]] = st["r"][o] + corner
    return out


def per_id(sim, key):
    st = O.compact(sim.state())
    out = np.zeros((len(st["id"]),) + st[key].shape[1:])
    out[st["id"]] = st[key]
    return out


# ---------------------------------------------------------------------------
# golden fixtures from the compiled reference

TRAJ = [
    ("two_atom_table.npz", "eam", False, "2.inp.gz"),
    ("two_atom_cheby.npz", "eam", True, "2.inp.gz"),
    ("cu6_table.npz", "eam", False, "cu6.inp.gz"),
    ("cu6_cheby.npz", "eam", True, "cu6.inp.gz"),
    ("cu6s_table.npz", "eam", False, "cu6s.inp.gz"),
    ("lj6.npz", "lj", False, "lj6.inp.gz"),
]


def make_pot(kind, cheby):
    if kind == "eam":
        p = O.OraclePot.eam(POTS, "cu500")
        p.set_eval(cheby)
        return p
    return O.OraclePot.lj(LJ_SIGMA, LJ_EPS, LJ_CUT, LJ_MASS)


@pytest.mark.parametrize("name,kind,cheby,inp", TRAJ, ids=[t[0][:-4] for t in TRAJ])
def test_oracle_reproduces_reference_trajectory(name, kind, cheby, inp):
    """Cell assignment, slot order and every per-atom value of the compiled
    reference after the initial force and after each nTimeSteps(K) call.
    Table mode is compared bit for bit; Chebyshev mode to 1e-12 relative,
    because its coefficients go through libm cos (src-flat/cheby.c:158,172),
    which is not correctly rounded and may differ between C libraries."""
    g = golden(name)
    pot = make_pot(kind, cheby)
    sim = O.OracleSim.gz(pot, data(inp))
    sim.rebox()
    sim.force()
    assert tuple(sim.info()["nbx"]) == tuple(g["nbx"])
    np.testing.assert_array_equal(sim.info()["boxsize"], g["boxsize"])
    for m in range(int(g["dumps"]) + 1):
        if m:
            sim.nsteps(int(g["steps"]))
        st = O.compact(sim.state())
        info = sim.info()
        np.testing.assert_array_equal(st["natoms"], g[f"s{m}_natoms"])
        np.testing.assert_array_equal(st["id"], g[f"s{m}_id"])
        for k in ("r", "p", "f", "rho", "fi"):
            if cheby:
                ref = g[f"s{m}_{k}"]
                scale = max(np.abs(ref).max(), 1e-300)
                assert np.abs(st[k] - ref).max() <= 1e-12 * scale, (m, k)
            else:
                np.testing.assert_array_equal(st[k], g[f"s{m}_{k}"], err_msg=f"state {m} {k}")
        if cheby:
            assert abs(info["e"] - float(g[f"s{m}_e"])) <= 1e-12 * abs(float(g[f"s{m}_e"]))
        else:
            assert info["e"] == float(g[f"s{m}_e"])
            assert info["stress"] == float(g[f"s{m}_stress"])


def test_oracle_reproduces_reference_silver_8k():
    """The reference's own data/8k.inp.gz with the silver tables."""
    g = golden("ag8k.npz")
    pot = O.OraclePot.eam(POTS, "ag")
    sim = O.OracleSim.gz(pot, data("8k.inp.gz"))
    sim.rebox()
    sim.force()
    st = O.compact(sim.state())
    np.testing.assert_array_equal(st["natoms"], g["s0_natoms"])
    np.testing.assert_array_equal(st["id"], g["s0_id"])
    for k in ("f", "rho", "fi"):
        np.testing.assert_array_equal(st[k], g[f"s0_{k}"])
    assert sim.info()["e"] == float(g["s0_e"])


def test_table_reader_matches_reference(cu):
    """Headers (x0 shifted by one interval, src-flat/eam.c:150), the values
    and guard slots (eam.c:232-246) as the reference read them."""
    g = golden("cu500_pot.npz")
    assert cu.cutoff == float(g["cutoff"]) and cu.mass == float(g["mass"]) and cu.lat == float(g["lat"])
    for k, name in enumerate(("phi", "rho", "F")):
        t = cu.table(k)
        n, x0, xn, invdx = g[f"{name}_hdr"]
        assert (t["n"], t["x0"], t["xn"], t["invdx"]) == (int(n), x0, xn, invdx)
        np.testing.assert_array_equal(t["values"][:4], g[f"{name}_head"])
        np.testing.assert_array_equal(t["values"][-4:], g[f"{name}_tail"])
    # decoded header values of cu500 (src-flat pots/cu500.*)
    phi, F = cu.table(0), cu.table(2)
    assert phi["n"] == 500 and abs(phi["invdx"] - 126.23075) < 1e-4
    assert abs(phi["x0"] - (1.0 + (4.961 - 1.0) / 500)) < 1e-12
    assert F["n"] == 2201 and abs(F["xn"] - 0.73916611) < 1e-8


def test_chebyshev_coefficients_match_reference(cu):
    g = golden("cu500_pot.npz")
    c = cu.cheb()
    ref = g["cheb"]
    assert c.shape == ref.shape == (6, 32)
    scale = np.abs(ref).max(axis=1, keepdims=True)
    assert (np.abs(c - ref) <= 1e-14 * scale).all()


def test_perfect_lattice_matches_reference():
    g = golden("fcc20.npz")
    for tag in ("table", "cheby"):
        pot = make_pot("eam", tag == "cheby")
        sim = O.OracleSim.fcc(pot, 20, 3.615)
        sim.rebox()
        sim.force()
        info = sim.info()
        assert tuple(info["nbx"]) == tuple(g["nbx"])
        assert abs(info["e"] - float(g[f"{tag}_e"])) <= 1e-13 * abs(float(g[f"{tag}_e"]))
        assert abs(info["stress"] - float(g[f"{tag}_stress"])) <= 1e-12 * abs(float(g[f"{tag}_stress"]))


# ---------------------------------------------------------------------------
# interpolation and Chebyshev closed forms


def test_interpolation_at_grid_points_and_clamps(cu):
    """At t = 0 the formula of src-flat/eam.c:401-434 gives v[i] and the
    central difference (v[i+1]-v[i-1])/2 * invDx; the derivative is zero in
    the first interval; arguments are clamped to [x0, xn]."""
    cu.set_eval(False)
    for which in (0, 1, 2):
        t = cu.table(which)
        v = t["values"][1:]  # v[0..n+1]
        dx = 1.0 / t["invdx"]
        for i in (1, 7, 50, t["n"] // 2, t["n"] - 2):
            # a hair above the grid point, so that rounding of (x-x0)*invDx
            # cannot drop it into interval i-1
            x = t["x0"] + (i + 1e-9) * dx
            val, der = cu.eval(which, x)
            assert abs(val - v[i]) <= 1e-8 * max(abs(v[i]), 1e-3)
            cd = 0.5 * (v[i + 1] - v[i - 1]) * t["invdx"]
            assert abs(der - cd) <= 1e-7 * max(abs(cd), 1e-3)
            # between grid points: the quadratic through (i-1, i, i+1)
            s = 0.37
            lag = v[i - 1] * s * (s - 1) / 2 + v[i] * (1 - s * s) + v[i + 1] * s * (s + 1) / 2
            val2, _ = cu.eval(which, t["x0"] + (i + s) * dx)
            assert abs(val2 - lag) <= 1e-10 * max(abs(lag), 1e-3)
        _, d0 = cu.eval(which, t["x0"] + 0.5 * dx)
        assert d0 == 0.0
        assert cu.eval(which, t["x0"] - 0.3) == cu.eval(which, t["x0"])
        assert cu.eval(which, t["xn"] + 0.3) == cu.eval(which, t["xn"])


def test_chebyshev_fit_interpolates_table_at_nodes(cu):
    """The DCT fit of src-flat/cheby.c:125-179 reproduces the sampled table
    at the m Chebyshev nodes (the defining property of the fit)."""
    m = 32
    for which in (0, 1, 2):
        t = cu.table(which)
        a, b = t["x0"], t["xn"]
        nodes = np.cos(np.pi * (np.arange(m) + 0.5) / m) * 0.5 * (b - a) + 0.5 * (b + a)
        cu.set_eval(False)
        tab = np.array([cu.eval(which, x)[0] for x in nodes])
        tab[tab > 1e10] = 0.0
        cu.set_eval(True)
        ser = np.array([cu.eval(which, x)[0] for x in nodes])
        assert np.abs(ser - tab).max() <= 1e-12 * np.abs(tab).max(), which
    cu.set_eval(False)


def test_chebyshev_clenshaw_and_derivative_series(cu):
    """Clenshaw (cheby.c:276-310) equals sum c_k T_k(y) - c_0/2; the
    derivative series (cheby.c:219-270) equals the derivative of the value
    series (central differences); only the lower end is clamped."""
    c = cu.cheb()
    cu.set_eval(True)
    rng = np.random.default_rng(3)
    for which in (0, 1, 2):
        t = cu.table(which)
        a, b = t["x0"], t["xn"]
        for x in rng.uniform(a, b, 12):
            y = (2 * x - a - b) / (b - a)
            direct = np.polynomial.chebyshev.chebval(y, c[which]) - 0.5 * c[which][0]
            val, der = cu.eval(which, x)
            assert abs(val - direct) <= 1e-13 * np.abs(c[which]).sum()
            h = 1e-5 * (b - a)
            fd = (cu.eval(which, x + h)[0] - cu.eval(which, x - h)[0]) / (2 * h)
            assert abs(der - fd) <= 1e-6 * max(np.abs(c[3 + which]).sum(), 1.0)
        assert cu.eval(which, a - 0.2) == cu.eval(which, a)
        # above b the series is evaluated, not clamped (cheby.c:293)
        yb = (2 * (b + 0.05) - a - b) / (b - a)
        above = np.polynomial.chebyshev.chebval(yb, c[which]) - 0.5 * c[which][0]
        assert abs(cu.eval(which, b + 0.05)[0] - above) <= 1e-12 * np.abs(c[which]).sum()
    cu.set_eval(False)


def test_chebyshev_derivative_basis_identity():
    """The derivative series (cheby.c:219-270) of basis function k,
    evaluated in the series convention (c_0 / 2), is
    (2 / (b - a)) k U_{k-1}(y): the identity by which the adjoint's list
    pass applies the derivative operator per pair instead of on the summed
    moments (csrc/adj_basis.h add_basis_deriv)."""
    from scipy.special import eval_chebyu

    pot = O.OraclePot.eam(POTS, "cu500")
    m = pot.cheb().shape[1]
    y = np.linspace(-1.0, 1.0, 41)
    for k in (1, 2, 5, 16, m - 1):
        coef = np.zeros((3, m))
        coef[:, k] = 1.0
        pot.set_cheb3(coef)
        d = pot.cheb()[3:6]
        for which in (0, 1, 2):
            t = pot.table(which)
            a, b = t["x0"], t["xn"]
            got = np.polynomial.chebyshev.chebval(y, d[which]) - 0.5 * d[which][0]
            want = 2.0 / (b - a) * k * eval_chebyu(k - 1, y)
            assert np.abs(got - want).max() <= 1e-12 * (2.0 / (b - a)) * k * k, (k, which)


# ---------------------------------------------------------------------------
# force closed forms and invariants


@pytest.mark.parametrize("cheby", [False, True])
def test_perfect_fcc_lattice_sum(cheby):
    """E/N = F(rho0) + 1/2 sum_s n_s phi(r_s), rho0 = sum_s n_s rho(r_s), over
    the shells (12, 6, 24) at a/sqrt2, a, a*sqrt(3/2) inside rc = 4.961; the
    4th shell (a*sqrt2 = 5.11) lies outside.  In table mode this is the
    cohesive energy of cu500 (3.54 eV); forces vanish by symmetry."""
    pot = make_pot("eam", cheby)
    a = 3.615
    shells = [(12, a / math.sqrt(2)), (6, a), (24, a * math.sqrt(1.5))]
    assert a * math.sqrt(2) > pot.cutoff
    rho0 = sum(n * pot.eval(1, r)[0] for n, r in shells)
    e_atom = pot.eval(2, rho0)[0] + 0.5 * sum(n * pot.eval(0, r)[0] for n, r in shells)
    sim = O.OracleSim.fcc(pot, 8, a)
    sim.rebox()
    sim.force()
    info = sim.info()
    assert abs(info["e"] / info["ntot"] - e_atom) <= 1e-12 * abs(e_atom)
    st = O.compact(sim.state())
    assert np.abs(st["f"][:, :3]).max() < 1e-12
    np.testing.assert_allclose(st["f"][:, 3], e_atom, rtol=1e-12)
    np.testing.assert_allclose(st["rho"], rho0, rtol=1e-12)
    if not cheby:
        assert abs(e_atom * HARTREE_EV + 3.5401) < 5e-4


@pytest.mark.parametrize("cheby", [False, True])
def test_two_atom_closed_form(cheby):
    """The reference's data/2.inp.gz: two atoms at minimum-image distance
    sqrt(4.29); E = phi(r) + 2 F(rho(r)) and f_1 = -f_2 = (phi' + 2 F'(rho)
    rho') dr/r (the EAM force array holds +grad U, src-flat/eam.c:529-533)."""
    pot = make_pot("eam", cheby)
    sim = O.OracleSim.gz(pot, data("2.inp.gz"))
    sim.rebox()
    sim.force()
    r = math.sqrt(4.29)
    phi, dphi = pot.eval(0, r)
    rho, drho = pot.eval(1, r)
    F, dF = pot.eval(2, rho)
    e = phi + 2 * F
    assert abs(sim.info()["e"] - e) <= 1e-13 * abs(e)
    f = per_id(sim, "f")
    pos = absolute_positions(sim)
    L = sim.info()["bounds"]
    d = pos[0] - pos[1]
    d -= L * np.round(d / L)
    assert abs(np.dot(d, d) - 4.29) < 1e-12
    g = (dphi + 2 * dF * drho) / r
    np.testing.assert_allclose(f[0, :3], g * d, rtol=1e-11, atol=1e-15)
    np.testing.assert_allclose(f[1, :3], -g * d, rtol=1e-11, atol=1e-15)
    np.testing.assert_allclose(per_id(sim, "rho"), [rho, rho], rtol=1e-13)


def test_lj_pair_closed_form():
    """U = 4 eps ((s/r)^12 - (s/r)^6); the LJ force array holds the physical
    force -grad U and the per-atom energies omit the factor 4 eps s^6
    (src-flat/ljForce.c:196-217)."""
    pot = O.OraclePot.lj(LJ_SIGMA, LJ_EPS, LJ_CUT, LJ_MASS)
    L = 3 * LJ_CUT
    r = 2.71
    xyz = np.array([[5.0, 6.0, 7.0], [5.0 + r * 0.6, 6.0 + r * 0.8, 7.0]])
    sim = O.OracleSim.from_atoms(pot, (L, L, L), xyz)
    sim.rebox()
    sim.force()
    sr6 = (LJ_SIGMA / r) ** 6
    U = 4 * LJ_EPS * (sr6 * sr6 - sr6)
    assert abs(sim.info()["e"] - U) <= 1e-13 * abs(U)
    f = per_id(sim, "f")
    dUdr = 4 * LJ_EPS * (-12 * sr6 * sr6 + 6 * sr6) / r
    u = (xyz[0] - xyz[1]) / r
    np.testing.assert_allclose(f[0, :3], -dUdr * u, rtol=1e-12, atol=1e-16)
    np.testing.assert_allclose(f[1, :3], dUdr * u, rtol=1e-12, atol=1e-16)
    np.testing.assert_allclose(f[:, 3].sum(), U / (4 * LJ_EPS * LJ_SIGMA ** 6), rtol=1e-12)


@pytest.mark.parametrize("kind,inp", [("eam", "cu6.inp.gz"), ("lj", "lj6.inp.gz")])
def test_newton_third_law_and_translation_invariance(kind, inp):
    pot = make_pot(kind, False)
    sim = O.OracleSim.gz(pot, data(inp))
    sim.rebox()
    sim.force()
    f = per_id(sim, "f")
    assert np.abs(f[:, :3].sum(0)).max() <= 1e-12 * np.abs(f[:, :3]).sum()
    e0 = sim.info()["e"]
    # the same atoms shifted by a vector and wrapped into the box
    pos = absolute_positions(sim)
    L = sim.info()["bounds"]
    shifted = np.mod(pos + np.array([0.31, 0.77, 1.13]), L)
    sim2 = O.OracleSim.from_atoms(pot, tuple(L), shifted)
    sim2.rebox()
    sim2.force()
    f2 = per_id(sim2, "f")
    scale = np.abs(f[:, :3]).max()
    assert np.abs(f2[:, :3] - f[:, :3]).max() <= 1e-10 * scale
    assert abs(sim2.info()["e"] - e0) <= 1e-12 * abs(e0)


# ---------------------------------------------------------------------------
# integrator


def test_free_flight():
    """With no neighbour inside the cutoff the force is zero and the DKD
    steps of src-flat/timestep.c:114-140 move the atom by n dt p / rmass."""
    pot = O.OraclePot.lj(LJ_SIGMA, LJ_EPS, LJ_CUT, LJ_MASS)
    L = 4 * LJ_CUT
    x0 = np.array([[1.0, 2.0, 3.0]])
    p0 = np.array([[0.9, -0.4, 0.25]]) * 1000.0
    sim = O.OracleSim.from_atoms(pot, (L, L, L), x0, p0)
    sim.rebox()
    sim.force()
    n = 40
    sim.nsteps(n)
    rmass = 1822.83 * LJ_MASS
    expect = np.mod(x0[0] + n * DT_INTERNAL * p0[0] / rmass, L)
    got = absolute_positions(sim)[0]
    assert np.abs(got - expect).max() < 1e-11
    assert (got // (L / sim.info()["nbx"][0])).tolist() != (x0[0] // (L / sim.info()["nbx"][0])).tolist()


def test_velocity_verlet_is_time_reversible():
    """n DKD steps, momenta negated, n more steps: back to the start."""
    pot = make_pot("eam", False)
    sim = O.OracleSim.gz(pot, data("cu6.inp.gz"))
    sim.rebox()
    sim.force()
    start = absolute_positions(sim)
    p_start = per_id(sim, "p")
    sim.nsteps(10)
    st = sim.state()
    sim.set_rp(st["r"], -st["p"])
    sim.force()
    sim.nsteps(10)
    L = sim.info()["bounds"]
    d = absolute_positions(sim) - start
    d -= L * np.round(d / L)
    assert np.abs(d).max() < 1e-9
    np.testing.assert_allclose(-per_id(sim, "p"), p_start, atol=1e-9 * np.abs(p_start).max())


def test_energy_error_is_second_order_in_dt():
    """KE + PE is conserved up to the Verlet error, which is second order:
    halving dt cuts the largest deviation over the same time span fourfold (a
    first-order slip in the kick/drift split would give 2).  Checked on the
    two-atom input in Chebyshev mode, where the force is the exact gradient
    of the energy and no pair crosses the cutoff (the tables' derivative is an
    interpolated central difference, not the derivative of the interpolated
    value, so there KE + PE also wanders at dt -> 0).  EAM f holds +grad U and
    the integrator subtracts it (src-flat/timestep.c:71)."""
    pot = make_pot("eam", True)
    dev = []
    for dt, n in ((1.0e-15, 40), (0.5e-15, 80)):
        sim = O.OracleSim.gz(pot, data("2.inp.gz"))
        sim.rebox()
        sim.force()
        e0 = sim.info()["e"] + sim.kinetic()
        worst = 0.0
        for _ in range(n):
            sim.nsteps(5, dt)
            worst = max(worst, abs(sim.info()["e"] + sim.kinetic() - e0))
        assert sim.kinetic() > 1e-3  # the dimer does oscillate
        dev.append(worst)
    assert 3.6 < dev[0] / dev[1] < 4.4, dev


# ---------------------------------------------------------------------------
# binning


def test_binning_floor_semantics_at_faces(cu):
    """Cell index floor(x / boxsize) per dimension (src-flat/domains.c:71-82),
    position kept relative to the lower cell corner; atoms exactly on a face
    belong to the upper cell, one ulp below it to the lower one."""
    L = 6 * 3.615
    sim0 = O.OracleSim.from_atoms(cu, (L, L, L), np.array([[1.0, 1.0, 1.0]]))
    nbx = sim0.info()["nbx"]
    bs = sim0.info()["boxsize"]
    pts = []
    for k in range(nbx[0]):
        face = k * bs[0]
        pts += [face, np.nextafter(face, -1.0) if k else 0.0, face + 0.5 * bs[0]]
    pts = np.array(pts)
    pts = pts[(pts >= 0) & (pts < L)]
    xyz = np.stack([pts, np.full_like(pts, 0.5 * bs[1]), np.full_like(pts, 2.5 * bs[2])], 1)
    sim = O.OracleSim.from_atoms(cu, (L, L, L), xyz)
    st = sim.state()
    for c in np.flatnonzero(st["natoms"]):
        for i in range(st["natoms"][c]):
            o = c * O.MAXATOMS + i
            x = xyz[st["id"][o]]
            want = [int(math.floor(x[j] / bs[j])) for j in range(3)]
            assert [c % nbx[0], (c // nbx[0]) % nbx[1], c // (nbx[0] * nbx[1])] == want
            corner = bs * np.array(want)
            np.testing.assert_array_equal(st["r"][o], x - corner)


def test_rebox_moves_and_swaps_with_last(cu):
    """reBoxAll (src-flat/domains.c:347-389): an atom whose relative x has
    gone negative moves to the lower neighbour (appended, r += boxsize) and
    the last atom of its cell takes its slot."""
    L = 6 * 3.615
    sim = O.OracleSim.from_atoms(cu, (L, L, L), np.array([[7.5, 7.5, 7.5], [7.6, 7.6, 7.6], [7.7, 7.7, 7.7]]))
    st = sim.state()
    c = np.flatnonzero(st["natoms"])[0]
    assert st["natoms"][c] == 3
    r = st["r"].copy()
    bs = sim.info()["boxsize"]
    r[c * 64 + 0, 0] = -1e-12  # first atom crosses the lower x face
    sim.set_rp(r, st["p"])
    sim.rebox()
    st2 = sim.state()
    assert st2["natoms"][c] == 2
    assert st2["id"][c * 64 + 0] == st["id"][c * 64 + 2]  # last one swapped in
    assert st2["id"][c * 64 + 1] == st["id"][c * 64 + 1]
    lower = c - 1
    assert st2["natoms"][lower] == 1 and st2["id"][lower * 64] == st["id"][c * 64]
    assert st2["r"][lower * 64, 0] == -1e-12 + bs[0]


def test_lj_reference_convention_is_unstable():
    """src-flat's LJ stores -grad U and its integrator subtracts it
    (ljForce.c, timestep.c:71), so atoms climb the potential: a 300 K FCC
    lattice gains kinetic energy without bound within about 50 fs (with
    copper's epsilon of 0.167 eV: x7 by step 40, x1e31 by step 50).  This
    is why trajectories that must stay physical use COMD_TUNE_LJ_PHYSICAL."""
    sigma, eps, mass = 2.315, 0.167 / 27.211386, 63.55  # epsilon in hartree
    olj = O.OraclePot.lj(sigma, eps, 2.5 * sigma, mass)
    lat = math.sqrt(2.0) * 2.0 ** (1.0 / 6.0) * sigma
    s = O.OracleSim.fcc(olj, 5, lat)
    st = s.state()
    rng = np.random.default_rng(2024)
    rmass = 1822.83 * mass
    p = rng.normal(0.0, math.sqrt(rmass * 3.166811563e-6 * 300.0), st["p"].shape)
    occupied = np.arange(p.shape[0]) % 64 < np.repeat(st["natoms"], 64)  # slot c*64+i used
    s.set_rp(st["r"], p * occupied[:, None])
    s.force()
    ke0 = s.kinetic()
    assert ke0 > 0.0
    s.nsteps(40)
    assert s.info()["err"] != 0 or s.kinetic() > 3.0 * ke0
    s.nsteps(10)
    assert s.info()["err"] != 0 or s.kinetic() > 1e20 * ke0


# ---------------------------------------------------------------------------
# virial tensor (orc_virial_tensor): pinned by the strain derivative of the
# oracle's own energy.  Under the homogeneous deformation x -> (I + e) x,
# dU/de_ab = sum_i (dU/dx_i,a) x_i,b = -sum_{i<j} f_ij,a r_ij,b, so the pair
# part of the virial is W_pair,ab = -dU/de_ba; the kinetic part has trace
# 2 KE.  A non-periodic cluster takes any strain (shear included); a
# periodic box takes the diagonal ones.  EAM is checked in Chebyshev form:
# its derivative series are the exact derivatives of its value series
# (src-flat/cheby.c:219-270), whereas the table form interpolates the
# derivative separately from the value (src-flat/eam.c:418-433), so that
# its forces differ from the gradient of its energy by ~1e-3 relative
# (test_table_forces_are_not_the_energy_gradient) and so would a strain
# derivative from the virial.

def _strained_energy(opot, xyz, bounds, eps, periodic, centre):
    x = centre + (xyz - centre) @ (np.eye(3) + eps).T
    b = tuple(bounds[k] * (1.0 + eps[k, k]) for k in range(3)) if periodic else bounds
    s = O.OracleSim.from_atoms(opot, b, x, None, periodic=periodic)
    assert s.info()["err"] == 0
    s.rebox()
    s.force()
    return s.info()["e"]


@pytest.mark.parametrize("kind,periodic", [("eam", False), ("eam", True), ("lj", False), ("lj", True)])
def test_oracle_virial_is_strain_derivative(kind, periodic):
    if kind == "eam":
        opot = O.OraclePot.eam(POTS, "cu500")
        opot.set_eval(True)
        lat, mass = 3.615, 63.55
    else:
        pot = CA.copper_lj()
        opot = O.OraclePot.lj(pot.sigma, pot.epsilon, pot.cutoff, pot.mass)
        lat, mass = pot.lat, pot.mass
    n = 4 if not periodic else 5
    xyz, p = CA.seed_state(CA.fcc_positions(n, n, n, lat), mass, 300.0, 0.1, 5)
    if periodic:
        bounds = (n * lat,) * 3
        centre = np.zeros(3)
    else:
        bounds = (50.0, 50.0, 50.0)
        xyz = xyz + (25.0 - 0.5 * n * lat)
        centre = np.full(3, 25.0)
    s = O.OracleSim.from_atoms(opot, bounds, xyz, p, periodic=periodic)
    assert min(s.info()["nbx"]) >= 3
    s.rebox()
    s.force()
    w, wpair = s.virial()
    st = s.state()
    m = (np.arange(64)[None, :] < st["natoms"][:, None]).reshape(-1)
    pp = st["p"][m]
    kin = pp.T @ pp / (CA.AMU_TO_ME * mass)
    np.testing.assert_allclose(w - wpair, kin, rtol=1e-12, atol=1e-12 * np.abs(kin).max())
    assert abs(np.trace(kin) - 2.0 * s.kinetic()) <= 1e-12 * s.kinetic()
    np.testing.assert_allclose(wpair, wpair.T, rtol=0, atol=1e-12 * np.abs(wpair).max())
    # two small steps: neither potential vanishes at its cutoff (the fitted
    # series and the unshifted LJ), so a pair that crosses it within a step
    # spoils that difference; one of the two must match
    comps = [(a, b) for a in range(3) for b in range(3)] if not periodic else [(0, 0), (1, 1), (2, 2)]
    for a, b in comps:
        fds = []
        for h in (1e-7, 3e-8):
            e = np.zeros((3, 3))
            e[a, b] = h
            up = _strained_energy(opot, xyz, bounds, e, periodic, centre)
            dn = _strained_energy(opot, xyz, bounds, -e, periodic, centre)
            fds.append(-(up - dn) / (2 * h))
        err = min(abs(fd - wpair[b, a]) for fd in fds)
        assert err <= 1e-6 * np.abs(wpair).max(), (kind, periodic, a, b, fds, wpair[b, a])


def test_table_forces_are_not_the_energy_gradient():
    """The reference's table interpolation (src-flat/eam.c:401-434) takes
    the derivative from a linear interpolation of central differences, not
    from the quadratic it uses for the value: its EAM forces differ from
    the gradient of its energy by about 1e-3 of the force (in Chebyshev form
    they agree to the finite difference's accuracy).  Both halves are
    pinned here; parity with the reference keeps its forces."""
    n, lat = 4, 3.615
    xyz, _ = CA.seed_state(CA.fcc_positions(n, n, n, lat), 63.55, 0.0, 0.1, 5)
    bounds = (50.0, 50.0, 50.0)
    xyz = xyz + (25.0 - 0.5 * n * lat)
    rel = {}
    for cheby in (False, True):
        opot = O.OraclePot.eam(POTS, "cu500")
        opot.set_eval(cheby)

        def energy(x):
            t = O.OracleSim.from_atoms(opot, bounds, x, None, periodic=False)
            t.rebox()
            t.force()
            return t.info()["e"]

        s = O.OracleSim.from_atoms(opot, bounds, xyz, None, periodic=False)
        s.rebox()
        s.force()
        st = s.state()
        m = (np.arange(64)[None, :] < st["natoms"][:, None]).reshape(-1)
        ids, f = st["id"][m], st["f"][m]
        errs = []
        for a in range(0, len(xyz), 37):
            k = np.flatnonzero(ids == a)[0]
            for c in range(3):
                xp, xm = xyz.copy(), xyz.copy()
                xp[a, c] += 1e-6
                xm[a, c] -= 1e-6
                errs.append(abs((energy(xp) - energy(xm)) / 2e-6 - f[k, c]) / np.abs(f[:, :3]).max())
        rel[cheby] = max(errs)
    assert rel[True] < 1e-6
    assert 1e-5 < rel[False] < 1e-2


# ---------------------------------------------------------------------------
# centrosymmetry (orc_centrosymmetry, src-viz/AtomVisualize.c:399-673)

def _matchings(items):
    """All perfect matchings of an even list, as lists of pairs (lowest first)."""
    if not items:
        return [[]]
    a, rest = items[0], items[1:]
    out = []
    for i, b in enumerate(rest):
        for m in _matchings(rest[:i] + rest[i + 1:]):
            out.append([(a, b)] + m)
    return out


_PAIRINGS = np.array(_matchings(list(range(12))))  # (10395, 6, 2)


def _brute_centro(xyz, bounds):
    """Minimum-image 12 nearest neighbours by a full sort, then the smallest
    sum over all pairings of |r_j + r_k|^2, in double."""
    L = np.asarray(bounds)
    out = np.empty(len(xyz))
    for i in range(len(xyz)):
        d = xyz - xyz[i]
        d -= L * np.round(d / L)
        r2 = (d ** 2).sum(1)
        r2[i] = np.inf
        nn = d[np.argsort(r2, kind="stable")[:12]]
        s = ((nn[:, None, :] + nn[None, :, :]) ** 2).sum(-1)
        out[i] = s[_PAIRINGS[:, :, 0], _PAIRINGS[:, :, 1]].sum(1).min()
    return out


def _centro_by_id(sim):
    st = sim.state()
    m = (np.arange(64)[None, :] < st["natoms"][:, None]).reshape(-1)
    c = np.empty(m.sum())
    c[st["id"][m]] = sim.centrosymmetry()[m]
    return c


def test_centrosymmetry_perfect_lattice_and_vacancy():
    """Perfect FCC: 0 for every atom.  A vacancy in a lattice displaced by
    0.02 A (so that the 12th neighbour of the vacancy's neighbours, one of
    six second-shell atoms at equal distance in the perfect lattice, is not
    a tie decided by scan order): its 12 neighbours stand out (about a^2/2
    and more), the others stay small, all as the brute-force search finds."""
    assert len(_PAIRINGS) == 10395
    pot = O.OraclePot.eam(POTS, "cu500")
    n, a = 5, 3.615
    xyz = CA.fcc_positions(n, n, n, a)
    bounds = (n * a,) * 3
    s = O.OracleSim.from_atoms(pot, bounds, xyz)
    assert np.abs(_centro_by_id(s)).max() < 1e-10
    gone = 137
    xd, _ = CA.seed_state(xyz, 63.55, 0.0, 0.02, 3)
    xv = np.delete(xd, gone, axis=0)
    s = O.OracleSim.from_atoms(pot, bounds, xv)
    c = _centro_by_id(s)
    ref = _brute_centro(xv, bounds)
    d = np.delete(xyz, gone, axis=0) - xyz[gone]
    d -= np.asarray(bounds) * np.round(d / np.asarray(bounds))
    near = (d ** 2).sum(1) < (0.75 * a) ** 2
    assert near.sum() == 12
    assert c[near].min() > 20 * c[~near].max()
    np.testing.assert_allclose(c, ref, rtol=1e-5, atol=2e-6)


def test_centrosymmetry_fewer_than_12_neighbours():
    """Atoms with fewer than 12 others in the 27 cells around them get 1000
    (src-viz/AtomVisualize.c:508): five atoms in a 20 A box."""
    pot = O.OraclePot.eam(POTS, "cu500")
    xyz = np.array([[1.0, 1.0, 1.0], [3.5, 1.0, 1.0], [1.0, 3.5, 1.0], [9.0, 9.0, 9.0], [15.0, 4.0, 12.0]])
    s = O.OracleSim.from_atoms(pot, (20.0, 20.0, 20.0), xyz)
    assert np.array_equal(_centro_by_id(s), np.full(5, 1000.0))


def test_centrosymmetry_matches_brute_force_thermal():
    pot = O.OraclePot.eam(POTS, "cu500")
    n, a = 5, 3.615
    xyz, _ = CA.seed_state(CA.fcc_positions(n, n, n, a), 63.55, 0.0, 0.15, 9)
    bounds = (n * a,) * 3
    s = O.OracleSim.from_atoms(pot, bounds, xyz)
    c = _centro_by_id(s)
    ref = _brute_centro(xyz, bounds)
    # the oracle sums each pairing in float and rounds image offsets to
    # float, as the reference does: about 1e-6 A per coordinate across the
    # box faces, 2 |r_j + r_k| 1e-6 in a parameter
    np.testing.assert_allclose(c, ref, rtol=1e-5, atol=2e-6)
    assert c.max() > 0.05
