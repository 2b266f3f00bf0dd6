"""Why the mixed-precision passes keep the pair geometry and the derivative
terms in FP64 (CPU only).

On a thermalised Cu state from the oracle (FCC 8^3, 300 K, 300 steps), the
per-atom force error against FP64 (|df| / max(|f|, rms |f|), the bar of the
mixed path: 1e-6) is computed for three ways of rounding the pair terms, all
with FP64 accumulation and the compact-row table formulas of the list
passes (src-flat/eam.c:418-433):

  * the library's mixed passes: separations exact (fixed-point differences),
    r^2, 1/r and the derivative terms in FP64, table values and F' in FP32;
  * the same with the derivative terms phi'(r) y and (F'_i + F'_j) rho'(r) y
    in FP32;
  * FP32 r^2 and 1/r as well.

The first stays well inside the bar (1e-7); the other two use up
most of it or exceed it on 2,048 atoms already (the maximum grows with the
atom count).  FP32 arithmetic issues at the FP64 rate on this device unless
it is packed, so the FP64 terms cost the mixed path no issue slots."""
import numpy as np

import comd_amd as CA
import oracle_py as O
from helpers import POTS, seeded_atoms


def thermal_state():
    xyz, p, bounds = seeded_atoms(8, 3.615, 63.55, temp=300.0, disp=0.0, seed=7)
    osim = O.OracleSim.from_atoms(O.OraclePot.eam(POTS, "cu500"), bounds, xyz, p)
    osim.rebox()
    osim.force()
    osim.nsteps(300)
    st, info = osim.state(), osim.info()
    bs, nbx = np.array(info["boxsize"]), np.array(info["nbx"])
    slots = np.flatnonzero((np.arange(64)[None, :] < st["natoms"][:, None]).reshape(-1))
    c = slots // 64
    cell = np.stack([c % nbx[0], (c // nbx[0]) % nbx[1], c // (nbx[0] * nbx[1])], 1)
    return cell * bs + st["r"][slots], st["f"][slots, :3], np.array(bounds)


def rows(tab):
    """Per-interval (d0, d1, c0, c1, c2) of the table formula."""
    v, n, inv = tab["values"], tab["n"], tab["invdx"]
    i = np.arange(n)
    vm, v0, v1, v2 = v[i], v[i + 1], v[i + 2], v[i + 3]  # v[0] holds values[-1]
    g0, g1 = v1 - vm, v2 - v0
    return np.stack([np.where(i <= 0, 0.0, 0.5 * g0 * inv), np.where(i <= 0, 0.0, 0.5 * (g1 - g0) * inv),
                     v0, 0.5 * g0, 0.5 * (v1 + vm - 2.0 * v0)], 1)


def interval(r, tab):
    u = np.maximum((r - tab["x0"]) * tab["invdx"], 0.0)
    i = np.minimum(np.floor(u).astype(np.int64), tab["n"] - 1)
    return i, u - i


def forces(x, L, pot, geo, der, val=np.float32):
    """EAM forces with r^2, 1/r in dtype geo, the derivative terms in der and
    the density values and F' in val."""
    tp, tr, tf = pot.tables
    P, R, F = rows(tp), rows(tr), rows(tf)
    I, J, D = [], [], []
    for a in range(0, len(x), 512):
        d = x[a:a + 512, None, :] - x[None, :, :]
        d -= L * np.rint(d / L)
        r2 = (d * d).sum(-1)
        ii, jj = np.nonzero((r2 <= pot.cutoff ** 2) & (r2 > 0.0))
        I.append(ii + a)
        J.append(jj)
        D.append(d[ii, jj])
    I, J, D = np.concatenate(I), np.concatenate(J), np.concatenate(D)
    r2 = (D * D).sum(1).astype(geo)
    y = (geo(1.0) / np.sqrt(r2)).astype(geo)
    i, t = interval((r2 * y).astype(np.float64), tp)
    s = ((P[i, 0] + P[i, 1] * t).astype(der) * y.astype(der)).astype(der)
    rho = np.zeros(len(x))
    np.add.at(rho, I, (R[i, 2] + t * (R[i, 3] + t * R[i, 4])).astype(val).astype(np.float64))
    iF, tF = interval(rho, tf)
    fp = (F[iF, 0] + F[iF, 1] * tF).astype(val).astype(der)
    g = ((fp[I] + fp[J]) * (R[i, 0] + R[i, 1] * t).astype(der) * y.astype(der)).astype(der)
    f = np.zeros((len(x), 3))
    for k in range(3):
        np.add.at(f[:, k], I, (s.astype(np.float64) + g.astype(np.float64)) * D[:, k])
    return f


def max_error(f, ref):
    mag = np.linalg.norm(ref, axis=1)
    return (np.linalg.norm(f - ref, axis=1) / np.maximum(mag, np.sqrt(np.mean(mag ** 2)))).max()


def test_fp64_derivative_terms_are_needed():
    pot = CA.EAMPotential("cu500")
    x, fref, L = thermal_state()
    # the FP64 restatement reproduces the oracle's forces (pins the formulas)
    f64 = forces(x, L, pot, np.float64, np.float64, np.float64)
    assert max_error(f64, fref) <= 1e-12
    mixed = max_error(forces(x, L, pot, np.float64, np.float64), f64)
    der32 = max_error(forces(x, L, pot, np.float64, np.float32), f64)
    geo32 = max_error(forces(x, L, pot, np.float32, np.float32), f64)
    print(f"[precision budget] max force error: FP64 derivative terms {mixed:.2e}, FP32 derivative terms "
          f"{der32:.2e}, FP32 r^2 and 1/r too {geo32:.2e}", flush=True)
    assert mixed <= 2e-7 < 5e-7 <= der32 < geo32
    assert geo32 > 1e-6
