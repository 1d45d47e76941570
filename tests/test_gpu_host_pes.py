"""A potential that lives on the host (a plugin with no device twin) driving the plain path
and the ring routines, against the oracle.

The "user's plugin" here is the coupled 2D double well written in numpy in this file, in the
role mcmod_waterdimer.f90:12-103 plays for the reference: the library calls its potforce
for the chain's beads at every kick (verletmodule.f90:573) and every U_M evaluation
(instantonmod.f90:105), and its Vdoubleprime (or central differences of potforce, step
1e-4, as mcmod_waterdimer.f90:83-103 does) for the ring Hessian (:184).  Bars: single steps
and U_M 1e-12 against the oracle; trajectories 1e-10; HVPs and log-determinants 1e-11 with
the analytic Hessian callback, 1e-6 with differenced gradients; the instanton action 1e-9
against the device potential's optimum.  The batched engine must refuse such a potential.
"""
import numpy as np
import pytest

import oracle as O
from helpers import oracle_pes, rand_chain, rel
from rptunnel import Batch, Chain, Pes, Ring, specs
from rptunnel._lib import RpError

pytestmark = pytest.mark.gpu

V0, A_, W, CC, M = 0.005, 0.7, 0.01, 0.3, 1836.0


def _potforce(x):
    """Coupled double well V0 (X^2/a^2 - 1)^2 + m w^2 (Y + c X^2)^2 / 2, x (2, nbead)."""
    X, Y = x
    t = X * X / A_ ** 2 - 1.0
    s = Y + CC * X * X
    e = V0 * t * t + 0.5 * M * W * W * s * s
    g = np.array([4.0 * V0 * t * X / A_ ** 2 + M * W * W * s * 2.0 * CC * X, M * W * W * s])
    return g, e


def _hessian(x):
    X, Y = x
    s = Y + CC * X * X
    h = np.zeros((x.shape[1], 2, 2))
    h[:, 0, 0] = 4.0 * V0 * (3.0 * X * X / A_ ** 2 - 1.0) / A_ ** 2 + M * W * W * (4.0 * CC * CC * X * X + 2.0 * CC * s)
    h[:, 0, 1] = h[:, 1, 0] = M * W * W * 2.0 * CC * X
    h[:, 1, 1] = M * W * W
    return h


@pytest.fixture(scope="module")
def setup():
    sp = specs.coupled2d(V0=V0, a=A_, w=W, c=CC, m=M)
    return sp, oracle_pes(sp), Pes.host(2, 1, _potforce, _hessian), Pes.host(2, 1, _potforce), Pes(sp)


def test_host_potential_per_bead_calls(setup):
    sp, op, host, host_fd, _ = setup
    rng = np.random.default_rng(1)
    x = rng.standard_normal((2, 7))
    g, e = host.potforce_beads(x)
    for b in range(7):
        go, eo = op.potforce(x[:, b])
        assert abs(e[b] - eo) <= 1e-12 * abs(eo) and rel(g[:, b], go) <= 1e-12
        assert rel(host.Vdoubleprime(x[:, b]), op.Vdoubleprime(x[:, b])) <= 1e-12
        assert rel(host_fd.Vdoubleprime(x[:, b]), op.Vdoubleprime(x[:, b])) <= 1e-6


def test_host_potential_plain_pile_and_andersen(setup):
    sp, op, host, _, dev = setup
    n, betan, tau, dt, gamma = 32, 2000.0 / 33, 20.0, 1.0, 1e-3
    x = rand_chain(sp, n, 3, scale=0.05)
    p = 2.0 * np.random.default_rng(4).standard_normal(x.shape)
    nm = O.NM(n, sp.ndim, sp.natom, sp.mass, betan, tau, dt, gamma, False, sp.well1, sp.well2)
    noise = np.random.default_rng(5).standard_normal(n * sp.ndof)
    ch = Chain(host, n, sp.mass, betan, tau, dt, gamma, sp.well1, sp.well2)
    ch.set_state(x, p)
    ch.time_step_pile(noise)
    xo, po = nm.time_step_pile(op, O.Rng(normals=noise), x, p)
    xg, pg = ch.get_state()
    assert rel(xg, xo) <= 1e-12 and rel(pg, po) <= 1e-12
    # a Langevin TI run and an Andersen run, from Philox streams, against the oracle
    dbdl = sp.well2 - sp.well1
    for thermostat in ("pile", "nm"):
        ch.set_state(x, p)
        ch.set_rng(31, 7)
        r = O.Rng(seed=31, stream=7)
        if thermostat == "pile":
            dh, _ = ch.propagate_pimd_pile(60, dbdl, imin=10)
            xo, po, dho, _ = O.propagate_pile(nm, op, r, x, p, 60, dbdl, imin=10)
        else:
            dh = ch.propagate_pimd_nm(60, dbdl, 5, imin=10)
            xo, po, dho = O.propagate_andersen(nm, op, r, x, p, 60, dbdl, 5, imin=10)
        xg, pg = ch.get_state()
        assert rel(xg, xo) <= 1e-10 and rel(pg, po) <= 1e-10, thermostat
        assert abs(dh - dho) <= 1e-10 * abs(dho), thermostat


def test_host_potential_ring_routines(setup):
    sp, op, host, host_fd, dev = setup
    ns, betan = [12, 40, 25], [15.0, 8.0, 30.0]
    xs = [rand_chain(sp, n, 10 + k, scale=0.05) for k, n in enumerate(ns)]
    host.set_V0(0.0)
    ring = Ring(host, sp.mass, ns, betan, sp.well1, sp.well2)
    ring_fd = Ring(host_fd, sp.mass, ns, betan, sp.well1, sp.well2)
    g, e = ring.UMforceenergy(xs)
    ring_fd.set_x(xs)
    rng = np.random.default_rng(12)
    vs = [rng.standard_normal((sp.ndof, n)) for n in ns]
    hv, hv_fd = ring.hvp(vs), ring_fd.hvp(vs)
    ld, nn, eta = ring.logdet(singlewell=False, skip_lowest=True)
    ld_fd, _, _ = ring_fd.logdet(singlewell=False, skip_lowest=True)
    for k, n in enumerate(ns):
        oc = O.Chain(n, sp.ndim, sp.natom, sp.mass, betan[k], True)
        go, eo = oc.UMforceenergy(op, xs[k], sp.well1, sp.well2)
        assert abs(e[k] - eo) <= 1e-12 * abs(eo) and rel(g[k], go) <= 1e-12
        band = oc.UMhessian(op, xs[k], False)
        assert rel(ring.UMhessian(k), band) <= 1e-11
        A = O.band_to_dense(band)
        # U_M's Hessian in the chain layout: the band's mass weighting undone, bead 1-2
        # coupling restored
        Hm = _um_hessian(sp, op, xs[k], betan[k])
        assert rel(hv[k].reshape(-1), Hm @ vs[k].reshape(-1)) <= 1e-11
        assert rel(hv_fd[k].reshape(-1), Hm @ vs[k].reshape(-1)) <= 1e-6
        ev = np.linalg.eigvalsh(A)
        assert nn[k] == np.sum(ev < 0)
        assert abs(ld[k] - O.lndetj(ev)) <= 1e-9 * abs(O.lndetj(ev))
        assert abs(ld_fd[k] - O.lndetj(ev)) <= 1e-6 * abs(O.lndetj(ev))
    # the instanton from a host potential is the device potential's instanton
    x0 = [np.outer(sp.well1, 1 - np.arange(n) / (n - 1)) + np.outer(sp.well2, np.arange(n) / (n - 1))
          for n in ns]
    _, act, _, st = ring.instanton(x0, pgtol=1e-9, factr=10.0)
    rdev = Ring(dev, sp.mass, ns, betan, sp.well1, sp.well2)
    _, actd, _, std = rdev.instanton(x0, pgtol=1e-9, factr=10.0)
    assert np.all(st == 0) and np.all(std == 0)
    assert np.all(np.abs(act - actd) <= 1e-9 * np.abs(actd)), (act, actd)


def _um_hessian(sp, op, x, betan):
    D, n = x.shape
    m = np.repeat(sp.mass, sp.ndim)
    H = np.zeros((n * D, n * D))
    for i in range(n):
        Hb = op.Vdoubleprime(x[:, i])
        for p in range(D):
            for q in range(D):
                H[p * n + i, q * n + i] += Hb[p, q]
    for p in range(D):
        k = m[p] / betan ** 2
        for i in range(n):
            H[p * n + i, p * n + i] += 2.0 * k
            if i + 1 < n:
                H[p * n + i, p * n + i + 1] -= k
                H[p * n + i + 1, p * n + i] -= k
    return H


def test_batched_engine_refuses_a_host_potential(setup):
    sp, _, host, _, _ = setup
    pol = dict(n=16, betan=10.0, xi=0.5, a=sp.well1, b=sp.well2, dbdl=sp.well2 - sp.well1, stream=0)
    with pytest.raises(RpError, match="plain path"):
        Batch(host, sp.mass, [pol], 20.0, 1.0, 1e-3)


def test_lowest_eigenvalues_by_bisection():
    """The head of detJ's ascending spectrum (instantonmod.f90:782-827) for chains through
    the EVB barrier with 0 to 9 negative eigenvalues: eigvalsh of the oracle's band."""
    sp = specs.evb()
    pes = Pes(sp)
    op = oracle_pes(sp)
    cases = [(8, 60.0), (12, 250.0), (24, 400.0)]
    ns = [n for n, _ in cases]
    betan = [b for _, b in cases]
    ring = Ring(pes, sp.mass, ns, betan, sp.well1, sp.well2)
    xs = [(np.outer(sp.well1, n - np.arange(1, n + 1)) + np.outer(sp.well2, np.arange(n))) / (n - 1)
          for n in ns]
    ring.set_x(xs)
    k = 12
    for singlewell in (False, True):
        got = ring.lowest_eigs(k, singlewell=singlewell)
        for c, n in enumerate(ns):
            oc = O.Chain(n, sp.ndim, sp.natom, sp.mass, betan[c], True)
            ev = O.detj_eigs(oc.UMhessian(op, xs[c], singlewell))[:k]
            scale = np.abs(ev).max()
            assert np.all(np.abs(got[c] - ev) <= 1e-10 * scale), (n, singlewell, got[c] - ev)
