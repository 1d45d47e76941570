"""Parity of the instanton kernels (action, adjoint gradient, HVP, ring Hessian, log-det,
L-BFGS optimum) against the oracle.

Bars: U_M and its gradient 1e-12 relative; HVPs and the UMhessian band 1e-11 against the
oracle's analytic potential Hessians (1e-6 for the water cluster, whose oracle Hessian is
central finite differences); log-determinants 1e-9 against numpy's eigvalsh of the
oracle's band; converged instanton actions 1e-9 relative against scipy's L-BFGS-B on the
oracle's U_M."""
import numpy as np
import pytest
import scipy.optimize as so

import oracle as O
from helpers import oracle_pes, rand_chain, rel
from rptunnel import Pes, Ring, specs

pytestmark = pytest.mark.gpu


def _chains(sp, ns, seed, scale=0.05):
    return [rand_chain(sp, n, seed + k, scale) for k, n in enumerate(ns)]


@pytest.mark.parametrize("name,fixed", [("evb", True), ("evb", False), ("quartic", True),
                                        ("water2", True), ("coupled2d", True)])
def test_um_and_gradient(name, fixed):
    sp = {"evb": specs.evb, "quartic": specs.quartic, "coupled2d": specs.coupled2d,
          "water2": lambda: specs.water_cluster(nmol=2)}[name]()
    pes = Pes(sp)
    pes.set_V0(0.001)
    op = oracle_pes(sp, 0.001)
    ns = [16, 37, 300]
    betan = [8.0, 12.5, 30.0]
    ring = Ring(pes, sp.mass, ns, betan, sp.well1, sp.well2, fixedends=fixed)
    xs = _chains(sp, ns, 3)
    g, e = ring.UMforceenergy(xs)
    for k, n in enumerate(ns):
        oc = O.Chain(n, sp.ndim, sp.natom, sp.mass, betan[k], fixed)
        go, eo = oc.UMforceenergy(op, xs[k], sp.well1, sp.well2)
        assert abs(e[k] - eo) <= 1e-12 * abs(eo)
        assert abs(oc.UM(op, xs[k], sp.well1, sp.well2) - eo) <= 1e-12 * abs(eo)
        assert rel(g[k], go) <= 1e-12
        assert rel(g[k], oc.UMprime(op, xs[k], sp.well1, sp.well2)) <= 1e-12


def _dense_hessian(oc, op, x):
    """Full (unweighted) Hessian of U_M from the oracle's bead Hessians and the springs."""
    n, D = x.shape[1], x.shape[0]
    H = np.zeros((n * D, n * D))
    m = np.repeat(oc.mass, oc.ndim)
    k = m / oc.s.betan ** 2
    for i in range(n):
        Hb = op.Vdoubleprime(x[:, i])  # [p, q] = d g_q / d x_p
        for p in range(D):
            for q in range(D):
                H[p * n + i, q * n + i] += Hb[p, q]
    for p in range(D):
        for i in range(n):
            H[p * n + i, p * n + i] += 2.0 * k[p]
            if i + 1 < n:
                H[p * n + i, p * n + i + 1] -= k[p]
                H[p * n + i + 1, p * n + i] -= k[p]
    return H


@pytest.mark.parametrize("name", ["evb", "quartic", "water2"])
def test_hvp(name):
    sp = {"evb": specs.evb, "quartic": specs.quartic,
          "water2": lambda: specs.water_cluster(nmol=2)}[name]()
    pes = Pes(sp)
    op = oracle_pes(sp)
    ns, betan = [12, 20], [10.0, 15.0]
    ring = Ring(pes, sp.mass, ns, betan, sp.well1, sp.well2)
    xs = _chains(sp, ns, 5)
    ring.set_x(xs)
    rng = np.random.default_rng(6)
    vs = [rng.standard_normal((sp.ndof, n)) for n in ns]
    hv = ring.hvp(vs)
    tol = 1e-6 if sp.kind == specs.PES_WATER else 1e-11
    for k, n in enumerate(ns):
        oc = O.Chain(n, sp.ndim, sp.natom, sp.mass, betan[k], True)
        H = _dense_hessian(oc, op, xs[k])
        assert rel(hv[k].reshape(-1), H @ vs[k].reshape(-1)) <= tol


@pytest.mark.parametrize("singlewell", [False, True])
def test_umhessian_band(singlewell):
    sp = specs.evb()
    pes = Pes(sp)
    op = oracle_pes(sp)
    n, betan = 10, 9.0
    ring = Ring(pes, sp.mass, [n], [betan], sp.well1, sp.well2)
    x = _chains(sp, [n], 7)
    ring.set_x(x)
    band = ring.UMhessian(0, singlewell)
    oc = O.Chain(n, sp.ndim, sp.natom, sp.mass, betan, True)
    bo = oc.UMhessian(op, x[0], singlewell)
    assert rel(band, bo) <= 1e-11
    # the reference's quirk: no coupling between beads 1 and 2 (instantonmod.f90:203)
    D = sp.ndof
    assert np.all(band[D, :D] == 0.0) and np.all(band[D, D:2 * D] == -1.0 / betan ** 2)


def test_logdet_at_well_against_eigvalsh():
    """lndetJ0 (rpi_ser.f90:221-236): every bead at well 1, singlewell, all eigenvalues."""
    sp = specs.evb()
    pes = Pes(sp)
    op = oracle_pes(sp)
    ns, betan = [8, 14], [60.0, 40.0]
    xs = [np.repeat(sp.well1[:, None], n, axis=1) for n in ns]
    ring = Ring(pes, sp.mass, ns, betan, sp.well1, sp.well2)
    ring.set_x(xs)
    ld0, nn0, _ = ring.logdet(singlewell=True, skip_lowest=False)
    for k, n in enumerate(ns):
        oc = O.Chain(n, sp.ndim, sp.natom, sp.mass, betan[k], True)
        ev = O.detj_eigs(oc.UMhessian(op, xs[k], True))
        assert nn0[k] == np.sum(ev < 0) == 0
        assert abs(ld0[k] - O.lndetj0(ev)) <= 1e-9 * abs(O.lndetj0(ev))


@pytest.mark.parametrize("D", [5, 54])
def test_lowest_eigen_and_logdet_harmonic_against_eigvalsh(D):
    """eta_1 by inverse iteration on the block factors, and lndetJ with and without it
    (rpi_ser.f90:221-236, 354-365), on rings of D = 5 and 54 dofs (the two lane layouts of
    the block solve): a harmonic surface, so the Hessian is that of the springs plus a
    diagonal, and eigvalsh of the dense band matrix is the check.  One soft heavy dof keeps
    eta_1 apart from eta_2 (inverse iteration cannot tell a near-degenerate pair apart)."""
    rng = np.random.default_rng(D)
    kf, mass = rng.uniform(0.1, 0.4, D), rng.uniform(500.0, 1000.0, D)
    kf[0], mass[0] = 0.02, 3000.0
    sp = specs.harmonic(k=kf, c=rng.normal(0.0, 0.3, D), mass=mass)
    pes = Pes(sp)
    op = oracle_pes(sp)
    ns, betan = [7, 33, 120], [25.0, 40.0, 12.0]
    xs = [sp.well1[:, None] + 0.1 * rng.standard_normal((D, n)) for n in ns]
    ring = Ring(pes, sp.mass, ns, betan, sp.well1, sp.well2)
    ring.set_x(xs)
    ld0, nn0, eta0 = ring.logdet(singlewell=False, skip_lowest=False, want_eta=True)
    ld, nn, eta = ring.logdet(singlewell=False, skip_lowest=True)
    for k, n in enumerate(ns):
        oc = O.Chain(n, sp.ndim, sp.natom, sp.mass, betan[k], True)
        ev = O.detj_eigs(oc.UMhessian(op, xs[k], False))
        assert nn0[k] == nn[k] == np.sum(ev < 0) == 0
        assert abs(eta0[k] - ev[0]) <= 1e-9 * abs(ev[0]) and eta[k] == eta0[k]
        assert abs(ld0[k] - O.lndetj0(ev)) <= 1e-9 * abs(O.lndetj0(ev))
        assert abs(ld[k] - O.lndetj(ev)) <= 1e-9 * abs(O.lndetj(ev))
    ring.close()


def _dense_host_pes(D, seed):
    """A host potential with dense D x D Hessians (no device twin has 28 <= D <= 39):
    V = (x - x0)^T K (x - x0) / 2 + 0.1 sum (x - x0)^4, K = QQ^T random, positive definite."""
    rng = np.random.default_rng(seed)
    Q = rng.standard_normal((D, D))
    K = 0.05 * (Q @ Q.T) / D + 0.01 * np.eye(D)
    x0 = rng.standard_normal(D)

    def potforce(x):
        d = x - x0[:, None]
        return K @ d + 0.4 * d ** 3, 0.5 * np.sum(d * (K @ d), axis=0) + 0.1 * np.sum(d ** 4, axis=0)

    def hessian(x):
        d = x - x0[:, None]
        h = np.repeat(K[None], x.shape[1], axis=0)
        h[:, np.arange(D), np.arange(D)] += 1.2 * d.T ** 2
        return h

    return Pes.host(3, D // 3, potforce, hessian), x0, np.full(D // 3, 1836.0)


@pytest.mark.parametrize("case", ["water2", "dense36", "water6"])
def test_logdet_dense_blocks_against_eigvalsh(case):
    """The block factorisation with dense D x D blocks (D = 18, 36 and 54 dofs, one per
    block-size class of the lower-triangle update: the 2- and 6-molecule water clusters on
    the device, a dense quartic host potential for 36) on chains near the minimum, against
    eigvalsh of the library's own band (rp_ring_hessian_band, the blocks the factorisation
    reads): the negative count, ln over the positive eigenvalues (rpi_ser.f90:227-236) 1e-9
    relative, and eta_1 when none is negative, 1e-8 relative."""
    rng = np.random.default_rng(len(case))
    if case == "dense36":
        pes, x0, mass = _dense_host_pes(36, 11)
        a = b = x0
        scale = 0.3
    else:
        sp = specs.water_cluster(nmol=int(case[-1]))
        pes, x0, mass, a, b = Pes(sp), sp.well1, sp.mass, sp.well1, sp.well2
        scale = 0.02
    D = pes.ndof
    ns, betan = [5, 23], [30.0, 18.0]
    xs = [x0[:, None] + scale * rng.standard_normal((D, n)) for n in ns]
    ring = Ring(pes, mass, ns, betan, a, b)
    ring.set_x(xs)
    ld0, nn0, eta0 = ring.logdet(singlewell=False, skip_lowest=False, want_eta=True)
    for k, n in enumerate(ns):
        ev = O.detj_eigs(ring.UMhessian(k, singlewell=False))
        assert nn0[k] == np.sum(ev < 0), (n, nn0[k], ev[:4])
        assert abs(ld0[k] - O.lndetj0(ev)) <= 1e-9 * abs(O.lndetj0(ev)), (n, ld0[k])
        if nn0[k] == 0:
            assert abs(eta0[k] - ev[0]) <= 1e-8 * ev[0], (n, eta0[k], ev[:3])
    ring.close()


def _linear_path(sp, n):
    k = np.arange(1, n + 1)
    return (np.outer(sp.well1, n - k) + np.outer(sp.well2, k - 1)) / (n - 1)


def test_instanton_converges_to_scipy_optimum():
    sp = specs.evb()
    pes = Pes(sp)
    op = oracle_pes(sp)
    ns = [16, 24]
    beta = [3000.0, 4000.0]
    betan = [b / n for b, n in zip(beta, ns)]  # rpi_ser.f90:45
    ring = Ring(pes, sp.mass, ns, betan, sp.well1, sp.well2)
    x0 = [_linear_path(sp, n) for n in ns]
    xo, act, it, st = ring.instanton(x0, pgtol=1e-9, factr=10.0)
    assert np.all(st == 0), st
    for k, n in enumerate(ns):
        oc = O.Chain(n, sp.ndim, sp.natom, sp.mass, betan[k], True)

        def fg(v):
            g, e = oc.UMforceenergy(op, v.reshape(sp.ndof, n), sp.well1, sp.well2)
            return e, g.reshape(-1)

        res = so.minimize(fg, x0[k].reshape(-1), jac=True, method="L-BFGS-B",
                          options={"maxcor": 8, "ftol": 1e-15, "gtol": 1e-10, "maxiter": 20000})
        assert abs(act[k] - res.fun) <= 1e-9 * abs(res.fun), (act[k], res.fun)
        # the device action at the returned chain agrees with the oracle's
        assert abs(ring.UM([xo[0], xo[1]])[k] - oc.UM(op, xo[k], sp.well1, sp.well2)) <= 1e-12 * abs(act[k])
    # lndetJ at the instanton (rpi_ser.f90:354-365): skip eta_1, sum ln over eta > 0
    ring.set_x(xo)
    ld, nn, eta = ring.logdet(singlewell=False, skip_lowest=True)
    for k, n in enumerate(ns):
        oc = O.Chain(n, sp.ndim, sp.natom, sp.mass, betan[k], True)
        ev = O.detj_eigs(oc.UMhessian(op, xo[k], False))
        assert nn[k] == np.sum(ev < 0)
        assert abs(eta[k] - ev[0]) <= 1e-8 * abs(ev[0]) + 1e-14
        assert abs(ld[k] - O.lndetj(ev)) <= 1e-9 * abs(O.lndetj(ev))


def test_instanton_batch_is_per_chain_and_caps():
    """Every chain of a batch follows its own search: chains of at most one bead tile give
    bit-identical optima alone and in a batch; maxiter caps the iterations (status 2); a
    start that already meets pgtol returns at once (status 0, no iteration)."""
    sp = specs.evb()
    pes = Pes(sp)
    ns = [16, 40, 12]
    betan = [3000.0 / n for n in ns]
    x0 = [_linear_path(sp, n) for n in ns]
    ring = Ring(pes, sp.mass, ns, betan, sp.well1, sp.well2)
    xo, act, it, st = ring.instanton(x0, pgtol=1e-8, factr=10.0)
    assert np.all(st == 0), st
    for k, n in enumerate(ns):
        one = Ring(pes, sp.mass, [n], [betan[k]], sp.well1, sp.well2)
        x1, a1, i1, s1 = one.instanton([x0[k]], pgtol=1e-8, factr=10.0)
        assert np.array_equal(x1[0], xo[k]) and a1[0] == act[k] and i1[0] == it[k]
        # the optimum stays on the device: U_M there is the returned action
        assert one.UM()[0] == pytest.approx(a1[0], rel=1e-14)
        one.close()
    xc, ac, ic, sc = ring.instanton(x0, maxiter=3)
    assert np.all(sc == 2) and np.all(ic == 3)
    assert np.all(ac < ring.UM(x0))
    xs, as_, is_, ss = ring.instanton(xo, pgtol=1.0)
    assert np.all(ss == 0) and np.all(is_ == 0)
    assert all(np.array_equal(a, b) for a, b in zip(xs, xo))


@pytest.mark.parametrize("nmol,ns", [(2, [9, 17]), (6, [128, 64])])
def test_water_param_grad_ring(nmol, ns):
    """dU_M/d(qH, De) of every chain: the bead sum of the oracle's sensitivities (pinned by
    central differences in the parameters), and central differences of the device's own
    U_M in each parameter (U_M is quadratic in qH and linear in De)."""
    sp = specs.water_cluster(nmol=nmol)
    pes = Pes(sp)
    op = oracle_pes(sp)
    ring = Ring(pes, sp.mass, ns, [10.0] * len(ns), sp.well1, sp.well2)
    xs = _chains(sp, ns, 9, scale=0.03)
    ring.set_x(xs)
    pg = ring.param_grad()
    for k, n in enumerate(ns):
        want = sum(op.param_grad(xs[k][:, i]) for i in range(n))
        assert rel(pg[k], want) <= 1e-12
    for j, idx, h in ((0, 7, 1e-3), (1, 0, 1e-3)):
        ums = []
        for sgn in (1.0, -1.0):
            dp = sp.dp.copy()
            dp[idx] += sgn * h
            spp = specs.PesSpec(sp.name, sp.kind, sp.ndim, sp.natom, dp, sp.ip, sp.mass,
                                sp.well1, sp.well2)
            r2 = Ring(Pes(spp), sp.mass, ns, [10.0] * len(ns), sp.well1, sp.well2)
            ums.append(r2.UM(xs))
            r2.close()
        fd = (ums[0] - ums[1]) / (2.0 * h)
        for k in range(len(ns)):
            assert pg[k][j] == pytest.approx(fd[k], rel=1e-8), (j, k, pg[k][j], fd[k])
