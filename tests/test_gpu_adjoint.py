"""GPU tests of the Chebyshev-coefficient adjoint (comd_eam_cheb_grad) and
its torch.autograd wrapper.

The gradient of the force-matching loss
    L(c) = sum_i |f_i(c) - t_i|^2 / (3N) + w_e (E(c) - E_t)^2
with respect to the 3 x 32 value coefficients of phi, rho and F is checked
against central finite differences with h = 1e-6 relative to each
coefficient (floored for coefficients near zero), computed on the CPU oracle
(an independent FP64 forward path) and on the GPU forward path.  Targets come
from the FP64 forward pass at coefficients perturbed by 1 % Gaussian noise.
"""
import numpy as np
import pytest

import comd_amd as CA
import oracle_py as O
from helpers import POTS, seeded_atoms

pytestmark = pytest.mark.gpu


def config(cu500, n, seed, precision="fp64"):
    xyz, p, bounds = seeded_atoms(n, 3.615, 63.55, temp=0.0, disp=0.1, seed=seed)
    sim = CA.Simulation(cu500, bounds, defgrad=0.0, precision=precision, evaluation="cheby")
    sim.bin_atoms(xyz, p)
    sim.rebox()
    return sim, xyz, bounds


def perturbed(c0, seed, rel=0.01):
    rng = np.random.default_rng(seed)
    return c0 * (1.0 + rel * rng.standard_normal(c0.shape))


def target_of(sim, coef):
    """FP64 forward at coef: slot-layout forces and the energy."""
    prec = sim.prm.precision
    sim.set_precision("fp64")
    sim.set_cheb(coef)
    sim.force()
    f = sim.download()["f"][:, :3].copy()
    e = sim.energy()[0]
    sim.set_precision("mixed" if prec == CA.PREC_MIXED else "fp64")
    return f, e


def steps(c0):
    """h_k = 1e-6 |c_k|, floored at 1e-9 of the largest coefficient of the
    same function (the high-order coefficients are tiny)."""
    floor = 1e-3 * np.abs(c0).max(axis=1, keepdims=True)
    return 1e-6 * np.maximum(np.abs(c0), floor)


def oracle_loss(osim, opot, coef, ft, et, w_e, mask):
    opot.set_cheb3(coef)
    osim.force()
    st = osim.state()
    d = st["f"][:, :3][mask] - ft[mask]
    n = mask.sum()
    return (d * d).sum() / (3 * n) + w_e * (osim.info()["e"] - et) ** 2


def compare(g, ref, what, tol):
    err = np.abs(g - ref).max(axis=1) / np.abs(ref).max(axis=1)
    print(f"[{what}] max |g - ref| / max|ref| per function (phi, rho, F): {err}", flush=True)
    assert (err <= tol).all(), f"{what}: {err}"


@pytest.mark.parametrize("w_e", [0.0, 1.0e-4])
def test_adjoint_matches_oracle_finite_differences(gpu, cu500, w_e):
    c0 = cu500.cheb[:3].copy()
    sim, xyz, bounds = config(cu500, 5, 11)
    ct = perturbed(c0, 5)
    ft, et = target_of(sim, ct)
    loss, g = sim.cheb_grad(c0, ft, et, w_e)
    st = sim.download()
    mask = (np.arange(64)[None, :] < st["natoms"][:, None]).reshape(-1)

    opot = O.OraclePot.eam(POTS, "cu500")
    opot.set_eval(True)
    osim = O.OracleSim.from_atoms(opot, bounds, xyz)
    osim.rebox()
    assert np.array_equal(O.compact(osim.state())["id"], st["id"][mask])
    l0 = oracle_loss(osim, opot, c0, ft, et, w_e, mask)
    print(f"[adjoint w_e={w_e}] loss gpu {loss:.12e} oracle {l0:.12e}", flush=True)
    assert abs(loss - l0) <= 1e-9 * abs(l0)
    h = steps(c0)
    fd = np.zeros_like(c0)
    for k in range(3):
        for j in range(c0.shape[1]):
            cp, cm = c0.copy(), c0.copy()
            cp[k, j] += h[k, j]
            cm[k, j] -= h[k, j]
            fd[k, j] = (oracle_loss(osim, opot, cp, ft, et, w_e, mask) -
                        oracle_loss(osim, opot, cm, ft, et, w_e, mask)) / (2 * h[k, j])
    compare(g, fd, f"adjoint vs oracle FD, w_e={w_e}", 1e-5)


def test_adjoint_matches_gpu_finite_differences_32k(gpu, cu500):
    """One configuration of the differentiation case: Cu FCC 20^3 (32 000
    atoms), displacements sigma 0.1 A, targets at 1 %-perturbed coefficients;
    finite differences of the GPU forward pass on 24 coefficients."""
    c0 = cu500.cheb[:3].copy()
    sim, _, _ = config(cu500, 20, 3)
    ft, et = target_of(sim, perturbed(c0, 9))
    w_e = 1.0 / 32000 ** 2
    loss, g = sim.cheb_grad(c0, ft, et, w_e)
    h = steps(c0)
    picks = [(k, j) for k in range(3) for j in (0, 1, 2, 3, 5, 8, 13, 21)]
    for k, j in picks:
        cp, cm = c0.copy(), c0.copy()
        cp[k, j] += h[k, j]
        cm[k, j] -= h[k, j]
        lp, _ = sim.cheb_grad(cp, ft, et, w_e)
        lm, _ = sim.cheb_grad(cm, ft, et, w_e)
        fd = (lp - lm) / (2 * h[k, j])
        scale = np.abs(g[k]).max()
        print(f"[32k] coef ({k},{j}) adjoint {g[k, j]:+.10e} fd {fd:+.10e}", flush=True)
        assert abs(g[k, j] - fd) <= 1e-5 * scale


def test_mixed_adjoint_against_fp64(gpu, cu500):
    c0 = cu500.cheb[:3].copy()
    sim, xyz, bounds = config(cu500, 6, 2)
    ft, et = target_of(sim, perturbed(c0, 4))
    l64, g64 = sim.cheb_grad(c0, ft, et, 1e-4)
    sim.set_precision("mixed")
    l32, g32 = sim.cheb_grad(c0, ft, et, 1e-4)
    rel = np.abs(g32 - g64).max(axis=1) / np.abs(g64).max(axis=1)
    print(f"[mixed adjoint] loss rel err {abs(l32 - l64) / l64:.3e}, gradient rel err per function {rel}", flush=True)
    assert abs(l32 - l64) <= 1e-4 * l64
    assert (rel <= 1e-3).all()


def test_adjoint_is_deterministic_and_set_cheb_round_trips(gpu, cu500):
    c0 = cu500.cheb[:3].copy()
    sim, _, _ = config(cu500, 6, 8)
    sim.force()
    f_ref = sim.download()["f"].copy()
    sim.set_cheb(c0)
    # the derivative series rebuilt by the library equal the host fit's
    np.testing.assert_array_equal(sim.get_cheb(), cu500.cheb)
    sim.force()
    np.testing.assert_array_equal(sim.download()["f"], f_ref)
    ft, et = target_of(sim, perturbed(c0, 1))
    a = sim.cheb_grad(c0, ft, et, 1e-4)
    b = sim.cheb_grad(c0, ft, et, 1e-4)
    assert a[0] == b[0]
    np.testing.assert_array_equal(a[1], b[1])


def test_autograd_function(gpu, cu500):
    torch = pytest.importorskip("torch")
    from comd_amd.autograd import force_matching_loss

    c0 = cu500.cheb[:3].copy()
    sims, targets = [], []
    for seed in (1, 2):
        sim, _, _ = config(cu500, 5, seed)
        targets.append(target_of(sim, perturbed(c0, 20 + seed)))
        sims.append(sim)
    coef = torch.tensor(c0.reshape(-1), dtype=torch.float64, requires_grad=True)
    loss = force_matching_loss(coef, sims, targets, 1e-4)
    loss.backward()
    expect = sum(s.cheb_grad(c0, t[0], t[1], 1e-4)[1] for s, t in zip(sims, targets))
    np.testing.assert_allclose(coef.grad.numpy().reshape(3, -1), expect, rtol=0, atol=0)
    # torch's own finite-difference check on a few coefficients
    idx = torch.tensor([0, 1, 33, 34, 64, 65])

    sub = coef.detach()[idx].clone().requires_grad_(True)
    assert torch.autograd.gradcheck(lambda s: force_matching_loss(
        coef.detach().index_put((idx,), s), sims, targets, 1e-4), (sub,), eps=1e-7, atol=1e-7, rtol=1e-3)


def test_batch_gradient_equals_sum_of_single_contexts(gpu, cu500):
    """comd_batch_cheb_grad over three configurations of different size in
    one batch (link-cell forward, one launch per adjoint stage) against the
    sum of comd_eam_cheb_grad over the same configurations in their own
    contexts: the loss is the sum of the losses (each with its own 1/(3N)
    and energy term) and the gradient the sum of the gradients."""
    c0 = cu500.cheb[:3].copy()
    w_e = 1e-4
    reps, fts, ets = [], [], []
    loss_sum, grad_sum = 0.0, np.zeros_like(c0)
    for k, (n, seed) in enumerate(((5, 31), (6, 32), (7, 33))):
        sim, xyz, bounds = config(cu500, n, seed)
        ft, et = target_of(sim, perturbed(c0, 40 + k))
        loss, g = sim.cheb_grad(c0, ft, et, w_e)
        loss_sum += loss
        grad_sum += g
        _, p, _ = seeded_atoms(n, 3.615, 63.55, temp=0.0, disp=0.1, seed=seed)
        reps.append((xyz, p, bounds))
        fts.append(ft)
        ets.append(et)
        sim.close()
    b = CA.Batch([dict(pot=cu500, bounds=bd, defgrad=0.0, evaluation="cheby") for _, _, bd in reps],
                 precision="fp64")
    for r, (xyz, p, _) in enumerate(reps):
        b.bin_atoms(r, xyz, p)
    b.rebox()
    # the same slot layout as the single contexts (the targets are per slot)
    for r in range(3):
        assert b.info(r)[1] * CA.MAXATOMS == fts[r].shape[0]
    loss_b, grad_b = b.cheb_grad(c0, fts, ets, w_e)
    print(f"[batch adjoint] loss {loss_b:.12e} sum of singles {loss_sum:.12e}", flush=True)
    assert abs(loss_b - loss_sum) <= 1e-10 * abs(loss_sum)
    compare(grad_b, grad_sum, "batch vs sum of singles", 1e-9)
    # reusing the batch's target
    loss_c, grad_c = b.cheb_grad(c0, None, ets, w_e)
    assert loss_c == loss_b
    np.testing.assert_array_equal(grad_c, grad_b)
    # the pair passes over the batch's lists (groups of every replica in one
    # launch) give the same
    assert b.list_info()["active"] == 1
    b.tune(adjoint_lists=2)
    loss_l, grad_l = b.cheb_grad(c0, None, ets, w_e)
    print(f"[batch adjoint over lists] loss {loss_l:.12e}", flush=True)
    assert abs(loss_l - loss_b) <= 1e-12 * abs(loss_b)
    compare(grad_l, grad_b, "batch adjoint, lists vs link cells", 1e-12)
    b.close()


# ---------------------------------------------------------------------------
# gradient of the energy alone (comd_eam_cheb_energy_grad)

def cheb_basis(x, dom, m):
    """dV/dc_k of the series value sum_{k>=1} c_k T_k(y) + c_0 / 2 at x
    (clamped below at the domain start, as the evaluation clamps it)."""
    a, b = dom
    y = (2.0 * np.maximum(x, a) - a - b) / (b - a)
    t = np.zeros((m,) + np.shape(x))
    t[0] = 1.0
    if m > 1:
        t[1] = y
    for k in range(2, m):
        t[k] = 2.0 * y * t[k - 1] - t[k - 2]
    t[0] = 0.5
    return t


def test_energy_gradient_closed_form_and_oracle_fd(gpu, cu500):
    """dE/dc at the fitted coefficients against (1) the closed forms
    dE/dc_phi = 1/2 sum_{i != j} B_k(r_ij), dE/dc_rho = sum_i F'(rho_i)
    sum_j B_k(r_ij), dE/dc_F = sum_i B_k(rho_i), from the pair distances of
    the atoms and the oracle's rho and F', and (2) central differences of
    the oracle's energy; the energy equals the oracle's."""
    c0 = cu500.cheb[:3].copy()
    m = c0.shape[1]
    sim, xyz, bounds = config(cu500, 5, 13)
    e, g = sim.cheb_energy_grad(c0)
    opot = O.OraclePot.eam(POTS, "cu500")
    opot.set_eval(True)
    osim = O.OracleSim.from_atoms(opot, bounds, xyz)
    osim.rebox()

    def oracle_energy(coef):
        opot.set_cheb3(coef)
        osim.force()
        return osim.info()["e"]

    e0 = oracle_energy(c0)
    print(f"[energy gradient] E gpu {e:.12e} oracle {e0:.12e}", flush=True)
    assert abs(e - e0) <= 1e-12 * abs(e0)
    # closed forms over ordered pairs inside the cutoff (minimum image)
    L = np.array(bounds)
    d = xyz[:, None, :] - xyz[None, :, :]
    d -= L * np.rint(d / L)
    r = np.sqrt((d * d).sum(-1))
    np.fill_diagonal(r, np.inf)
    inside = r * r <= cu500.cutoff ** 2
    ii, jj = np.nonzero(inside)
    bphi = cheb_basis(r[ii, jj], cu500.cheb_dom[0], m)
    brho = cheb_basis(r[ii, jj], cu500.cheb_dom[1], m)
    st = osim.state()
    occ = (np.arange(64)[None, :] < st["natoms"][:, None]).reshape(-1)
    rho = np.zeros(len(xyz))
    fp = np.zeros(len(xyz))
    rho[st["id"][occ]] = st["rho"][occ]
    fp[st["id"][occ]] = st["fi"][occ]
    closed = np.stack([0.5 * bphi.sum(1), (brho * fp[ii]).sum(1), cheb_basis(rho, cu500.cheb_dom[2], m).sum(1)])
    compare(g, closed, "dE/dc vs closed form", 1e-10)
    h = steps(c0)
    fd = np.zeros_like(c0)
    for k in range(3):
        for j in range(m):
            cp, cm = c0.copy(), c0.copy()
            cp[k, j] += h[k, j]
            cm[k, j] -= h[k, j]
            fd[k, j] = (oracle_energy(cp) - oracle_energy(cm)) / (2 * h[k, j])
    compare(g, fd, "dE/dc vs oracle FD", 2e-6)  # the differences' rounding: ~1e-16 |E| / h


def test_batch_energy_gradient_is_weighted_sum(gpu, cu500):
    """comd_batch_cheb_energy_grad over three configurations: the weighted
    sum of the single contexts' energy gradients, and each E_r theirs; the
    mixed path (Chebyshev basis in FP32, sums in FP64) within 2e-5 of FP64
    relative to each function's largest component."""
    c0 = cu500.cheb[:3].copy()
    specs = [(4, 21), (5, 22), (6, 23)]
    w = np.array([1.0, -0.5, 2.0])
    reps, singles = [], []
    for n, seed in specs:
        sim, xyz, bounds = config(cu500, n, seed)
        reps.append((xyz, bounds))
        singles.append(sim.cheb_energy_grad(c0))
    b = CA.Batch([dict(pot=cu500, bounds=bd, defgrad=0.0, evaluation="cheby") for _, bd in reps], precision="fp64")
    for r, (xyz, _) in enumerate(reps):
        b.bin_atoms(r, xyz, None)
    b.rebox()
    e, g = b.cheb_energy_grad(c0, w)
    expect = sum(wk * gs for wk, (_, gs) in zip(w, singles))
    for r, (es, _) in enumerate(singles):
        assert abs(e[r] - es) <= 1e-12 * abs(es)
    compare(g, expect, "batch dE/dc vs weighted singles", 1e-12)
    sim, _, _ = config(cu500, 6, 23, precision="mixed")
    e32, g32 = sim.cheb_energy_grad(c0)
    compare(g32, singles[2][1], "mixed dE/dc vs FP64", 2e-5)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_adjoint_over_lists_equals_link_cells(gpu, cu500, precision):
    """The adjoint's pair passes over the neighbour lists (COMD_TUNE_ADJOINT_LISTS
    2; by default they run where the context has a list group per CU) and
    over the link cells (0) give the same loss and gradient, and the same
    energy gradient: FP64 to 1e-12 of each function's largest component,
    mixed (FP32 basis values of separations rounded differently) to 1e-6;
    both deterministic."""
    import time
    c0 = cu500.cheb[:3].copy()
    sim, xyz, bounds = config(cu500, 10, 17, precision)
    sim.force()
    assert sim.list_info()["active"] == 1
    ft, et = target_of(sim, perturbed(c0, 3))
    out = {}
    for lists in (2, 0):
        sim.tune(adjoint_lists=lists)
        sim.cheb_grad(c0, ft, et, 1e-4)  # warm-up
        t0 = time.perf_counter()
        l, g = sim.cheb_grad(c0, None, et, 1e-4)
        dt = time.perf_counter() - t0
        l2, g2 = sim.cheb_grad(c0, None, et, 1e-4)
        assert l2 == l and np.array_equal(g2, g)
        e, ge = sim.cheb_energy_grad(c0)
        out[lists] = (l, g, e, ge)
        print(f"[adjoint {precision} lists={lists}] {1e3 * dt:.2f} ms per gradient, loss {l:.12e}", flush=True)
    tol = 1e-12 if precision == "fp64" else 1e-6
    (l1, g1, e1, ge1), (l0, g0, e0, ge0) = out[2], out[0]
    assert abs(l1 - l0) <= tol * abs(l0)
    assert abs(e1 - e0) <= 1e-12 * abs(e0)
    compare(g1, g0, f"{precision} loss gradient, lists vs link cells", tol)
    compare(ge1, ge0, f"{precision} energy gradient, lists vs link cells", tol)


def test_batch_16x32k_gradient_at_size(gpu, cu500):
    """The differentiation case at its size: 16 configurations of Cu FCC
    20^3 (32 000 atoms each, displacements sigma 0.1 A) in one batch, whose
    adjoint runs over the neighbour lists; targets at 1 %-perturbed
    coefficients.  The FP64 gradient against central differences of the
    batch's own loss on 12 coefficients, and the mixed gradient against the
    FP64 one per function."""
    c0 = cu500.cheb[:3].copy()
    nconf, n = 16, 20
    reps, states = [], []
    for k in range(nconf):
        xyz, p, bounds = seeded_atoms(n, 3.615, 63.55, temp=0.0, disp=0.1, seed=500 + k)
        reps.append(dict(pot=cu500, bounds=bounds, defgrad=0.0, evaluation="cheby"))
        states.append((xyz, p))
    b = CA.Batch(reps, precision="fp64")
    for r, (xyz, p) in enumerate(states):
        b.bin_atoms(r, xyz, p)
    b.rebox()
    b.set_cheb(perturbed(c0, 77))
    b.force()
    fts = [b.download(r)["f"][:, :3].copy() for r in range(nconf)]
    ets = [b.energy_of(r)[0] for r in range(nconf)]
    w_e = 1.0 / (4 * n ** 3) ** 2
    loss, g = b.cheb_grad(c0, fts, ets, w_e)
    assert b.list_info()["active"] == 1
    h = steps(c0)
    for k, j in [(k, j) for k in range(3) for j in (0, 1, 4, 11)]:
        cp, cm = c0.copy(), c0.copy()
        cp[k, j] += h[k, j]
        cm[k, j] -= h[k, j]
        lp, _ = b.cheb_grad(cp, None, ets, w_e)
        lm, _ = b.cheb_grad(cm, None, ets, w_e)
        fd = (lp - lm) / (2 * h[k, j])
        print(f"[16x32k] coef ({k},{j}) adjoint {g[k, j]:+.10e} fd {fd:+.10e}", flush=True)
        assert abs(g[k, j] - fd) <= 1e-5 * np.abs(g[k]).max()
    b.set_precision("mixed")
    lm32, g32 = b.cheb_grad(c0, None, ets, w_e)
    rel = np.abs(g32 - g).max(axis=1) / np.abs(g).max(axis=1)
    print(f"[16x32k] mixed: loss rel err {abs(lm32 - loss) / loss:.2e}, gradient rel err per function {rel}",
          flush=True)
    assert abs(lm32 - loss) <= 1e-5 * loss
    assert (rel <= 1e-4).all()
    b.close()
