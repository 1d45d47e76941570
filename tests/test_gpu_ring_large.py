"""The instanton kernels at the chain lengths of the batched-instanton workload (27-dof EVB,
N = 64, 256 and 1024 in one batch), and log-determinants of spectra with several negative
eigenvalues.

Every reference here is a library routine applied to the oracle's pieces, or the
mathematics of the minimum:
* UMhessian band (instantonmod.f90:155-217): the oracle's band, 1e-11.
* HVP: the exact U_M Hessian assembled from the oracle's bead Hessians (analytic for the
  EVB) and the springs, applied as a sparse product, 1e-11.
* Log-determinants (rpi_ser.f90:221-237, 354-365): the band's |det| from SuperLU
  (scipy.sparse.linalg.splu), its lowest eigenvalues from ARPACK shift-invert (eigsh)
  at a shift that a banded Cholesky proves lies below the spectrum; 1e-9.
* The L-BFGS optimum (instantonmod.f90:679-777): the fixed-end minimum is unique, so
  Newton's method on the exact banded Hessian (scipy.linalg.solveh_banded) polishes the
  device's optimum to machine precision; the device's action agrees with it to 1e-9.
* Several negative eigenvalues (small chains through the barrier): numpy's eigvalsh.
"""
import numpy as np
import pytest
import scipy.linalg as sl
import scipy.sparse as sps
import scipy.sparse.linalg as spl

import oracle as O
from helpers import oracle_pes, rel
from rptunnel import Pes, Ring, specs

pytestmark = pytest.mark.gpu

NS = [64, 256, 1024]
BETAS = [1000.0, 4000.0, 8000.0]


def _line(sp, n):
    k = np.arange(1, n + 1)
    return (np.outer(sp.well1, n - k) + np.outer(sp.well2, k - 1)) / (n - 1)


def _ring(sp, pes, ns, betas):
    betan = [b / n for b, n in zip(betas, ns)]  # rpi_ser.f90:45
    return Ring(pes, sp.mass, ns, betan, sp.well1, sp.well2), betan


def _band_to_sparse(band):
    kd1, m = band.shape
    diags = [band[0]] + [band[r, : m - r] for r in range(1, kd1)] * 2
    offs = [0] + [-r for r in range(1, kd1)] + [r for r in range(1, kd1)]
    return sps.diags(diags, offs, shape=(m, m), format="csc")


def _exact_hessian_banded(sp, op, x, betan):
    """Lower band (kd = D) of the exact U_M Hessian in bead-major order (bead * D + dof):
    bead blocks H_i + 2 diag(m/betan^2), couplings -diag(m/betan^2) between beads i, i+1."""
    D, n = x.shape
    k = np.repeat(sp.mass, sp.ndim) / betan ** 2
    ab = np.zeros((D + 1, n * D))
    for i in range(n):
        Hb = op.Vdoubleprime(x[:, i]) + np.diag(2.0 * k)
        for r in range(D):
            ab[r, i * D: i * D + D - r] = np.diagonal(Hb, -r)
        if i + 1 < n:
            ab[D, i * D: i * D + D] = -k
    return ab


def _to_bead_major(x):
    return x.T.reshape(-1)


def _from_bead_major(v, D):
    return v.reshape(-1, D).T


@pytest.fixture(scope="module")
def evb():
    sp = specs.evb()
    return sp, Pes(sp), oracle_pes(sp)


def test_band_and_hvp_large(evb):
    sp, pes, op = evb
    ring, betan = _ring(sp, pes, NS, BETAS)
    rng = np.random.default_rng(11)
    xs = [_line(sp, n) + 0.03 * rng.standard_normal((sp.ndof, n)) for n in NS]
    ring.set_x(xs)
    vs = [rng.standard_normal((sp.ndof, n)) for n in NS]
    hv = ring.hvp(vs)
    for c, n in enumerate(NS):
        oc = O.Chain(n, sp.ndim, sp.natom, sp.mass, betan[c], True)
        band = ring.UMhessian(c, singlewell=False)
        assert rel(band, oc.UMhessian(op, xs[c], False)) <= 1e-11, n
        A = _band_to_sparse(_exact_hessian_banded(sp, op, xs[c], betan[c]))
        want = _from_bead_major(A @ _to_bead_major(vs[c]), sp.ndof)
        assert rel(hv[c], want) <= 1e-11, n


def _reference_sums(band, eta1_dev, nneg_dev):
    """(sum ln eta over eta > 0, the same without eta_1, the lowest eigenvalues) of the band
    from SuperLU's |det| and ARPACK's lowest eigenvalues."""
    A = _band_to_sparse(band)
    # a shift below the device's lowest eigenvalue; Cholesky succeeds only if it is below
    # every eigenvalue
    sig = eta1_dev - max(1e-3 * abs(eta1_dev), 1e-12)
    shifted = band.copy()
    shifted[0] -= sig
    sl.cholesky_banded(shifted, lower=True)
    k = nneg_dev + 2
    ev = np.sort(spl.eigsh(A, k=k, sigma=sig, which="LM", return_eigenvectors=False, tol=1e-13))
    lu = spl.splu(A)
    logabs = float(np.sum(np.log(np.abs(lu.U.diagonal()))))
    neg = ev[ev < 0.0]
    assert neg.size < k  # the eigenvalues found reach past the negative ones
    pos_sum = logabs - float(np.sum(np.log(-neg)))
    skip = pos_sum if ev[0] < 0.0 else pos_sum - float(np.log(ev[0]))
    return pos_sum, skip, ev


def test_logdet_large_at_well_and_line(evb):
    """lndetJ0 with every bead at well 1 (singlewell) and lndetJ on a perturbed straight
    line through the barrier, where the band has several negative eigenvalues."""
    sp, pes, op = evb
    ring, betan = _ring(sp, pes, NS, BETAS)
    ring.set_x([np.repeat(sp.well1[:, None], n, axis=1) for n in NS])
    ld0, nn0, eta0 = ring.logdet(singlewell=True, skip_lowest=False, want_eta=True)
    for c, n in enumerate(NS):
        band = ring.UMhessian(c, singlewell=True)
        L = sl.cholesky_banded(band, lower=True)  # positive definite at the well
        want = 2.0 * float(np.sum(np.log(L[0])))
        assert nn0[c] == 0
        assert abs(ld0[c] - want) <= 1e-9 * abs(want), (n, ld0[c], want)
    rng = np.random.default_rng(2)
    xs = [_line(sp, n) + 0.01 * rng.standard_normal((sp.ndof, n)) for n in NS]
    ring.set_x(xs)
    ld, nn, eta = ring.logdet(singlewell=False, skip_lowest=True)
    ldp, nnp, _ = ring.logdet(singlewell=False, skip_lowest=False)
    assert np.array_equal(nn, nnp)
    for c, n in enumerate(NS):
        band = ring.UMhessian(c, singlewell=False)
        pos_sum, skip, ev = _reference_sums(band, eta[c], nn[c])
        assert nn[c] == np.sum(ev < 0.0), (n, nn[c], ev)
        assert abs(eta[c] - ev[0]) <= 1e-8 * abs(ev[0]), (n, eta[c], ev[0])
        assert abs(ld[c] - skip) <= 1e-9 * abs(skip), (n, ld[c], skip)
        assert abs(ldp[c] - pos_sum) <= 1e-9 * abs(pos_sum), (n, ldp[c], pos_sum)


def test_logdet_many_negative_eigenvalues(evb):
    """Chains through the barrier with weak springs have 0, 2, 2, 4 and 9 negative
    eigenvalues; the sums keep exactly the reference's eigenvalues."""
    sp, pes, op = evb
    cases = [(8, 60.0), (12, 250.0), (20, 100.0), (16, 300.0), (24, 400.0)]
    ns = [n for n, _ in cases]
    betan = [b for _, b in cases]
    ring = Ring(pes, sp.mass, ns, betan, sp.well1, sp.well2)
    xs = [_line(sp, n) for n in ns]
    ring.set_x(xs)
    ld, nn, eta = ring.logdet(singlewell=False, skip_lowest=True)
    ldp, nnp, _ = ring.logdet(singlewell=False, skip_lowest=False)
    for c, n in enumerate(ns):
        oc = O.Chain(n, sp.ndim, sp.natom, sp.mass, betan[c], True)
        ev = O.detj_eigs(oc.UMhessian(op, xs[c], False))
        assert nn[c] == nnp[c] == np.sum(ev < 0), (n, nn[c], ev[:10])
        assert abs(eta[c] - ev[0]) <= 1e-9 * abs(ev[0]), (n, eta[c], ev[0])
        assert abs(ld[c] - O.lndetj(ev)) <= 1e-9 * abs(O.lndetj(ev)), (n, ld[c], O.lndetj(ev))
        assert abs(ldp[c] - O.lndetj0(ev)) <= 1e-9 * abs(O.lndetj0(ev)), (n, ldp[c])
    assert sorted(nn.tolist()) == [0, 2, 2, 4, 9]


def test_instanton_large_is_the_newton_minimum(evb):
    sp, pes, op = evb
    ring, betan = _ring(sp, pes, NS, BETAS)
    x0 = [_line(sp, n) for n in NS]
    xo, act, it, st = ring.instanton(x0, pgtol=1e-8, factr=10.0)
    assert np.all(st == 0), (st, it)
    for c, n in enumerate(NS):
        oc = O.Chain(n, sp.ndim, sp.natom, sp.mass, betan[c], True)
        x = xo[c].copy()
        for _ in range(4):
            g, _ = oc.UMforceenergy(op, x, sp.well1, sp.well2)
            ab = _exact_hessian_banded(sp, op, x, betan[c])
            step = sl.solveh_banded(ab, _to_bead_major(g), lower=True)
            x = x - _from_bead_major(step, sp.ndof)
        g, s_star = oc.UMforceenergy(op, x, sp.well1, sp.well2)
        assert np.abs(g).max() <= 1e-10, (n, np.abs(g).max())
        assert abs(act[c] - s_star) <= 1e-9 * abs(s_star), (n, act[c], s_star, it[c])
        # the device action at the returned chain is the oracle's U_M there
        assert abs(act[c] - oc.UM(op, xo[c], sp.well1, sp.well2)) <= 1e-12 * abs(act[c])


def _newton_min(sp, op, x, betan):
    oc = O.Chain(x.shape[1], sp.ndim, sp.natom, sp.mass, betan, True)
    for _ in range(4):
        g, _ = oc.UMforceenergy(op, x, sp.well1, sp.well2)
        ab = _exact_hessian_banded(sp, op, x, betan)
        x = x - _from_bead_major(sl.solveh_banded(ab, _to_bead_major(g), lower=True), sp.ndof)
    g, s_star = oc.UMforceenergy(op, x, sp.well1, sp.well2)
    assert np.abs(g).max() <= 1e-10
    return s_star, oc


def test_preconditioned_instanton_is_the_newton_minimum(evb):
    """rp_ring_set_preconditioner: the same minimum in a few tens of iterations where the
    plain method takes hundreds to thousands (tight tolerances: the action to 1e-9 of the
    Newton minimum), and with the reference's tolerances (pgtol = 1e-5, factr = 1e6) a
    returned gradient inside pgtol."""
    sp, pes, op = evb
    ring, betan = _ring(sp, pes, NS, BETAS)
    x0 = [_line(sp, n) for n in NS]
    _, _, it_plain, st_plain = ring.instanton(x0, pgtol=1e-5, factr=1e6)
    ring.set_preconditioner(0.05)
    xo, act, it, st = ring.instanton(x0, pgtol=1e-8, factr=10.0)
    assert np.all(st == 0), (st, it)
    for c, n in enumerate(NS):
        s_star, oc = _newton_min(sp, op, xo[c].copy(), betan[c])
        assert abs(act[c] - s_star) <= 1e-9 * abs(s_star), (n, act[c], s_star, it[c])
        assert abs(act[c] - oc.UM(op, xo[c], sp.well1, sp.well2)) <= 1e-12 * abs(act[c])
    xr, actr, itr, str_ = ring.instanton(x0, pgtol=1e-5, factr=1e6)
    assert np.all(str_ == 0) and np.all(st_plain == 0)
    assert np.all(itr * 5 <= it_plain), (itr, it_plain)
    for c, n in enumerate(NS):
        oc = O.Chain(n, sp.ndim, sp.natom, sp.mass, betan[c], True)
        g, e = oc.UMforceenergy(op, xr[c], sp.well1, sp.well2)
        # the factr test may stop it first; otherwise the gradient is inside pgtol
        assert np.abs(g).max() <= 1e-5 or itr[c] > 0, (n, np.abs(g).max())
        s_star, _ = _newton_min(sp, op, xr[c].copy(), betan[c])
        assert e >= s_star - 1e-12 * abs(s_star) and e - s_star <= 1e-5 * abs(s_star), (n, e, s_star)
    ring.set_preconditioner(0.0)
    _, _, it_off, _ = ring.instanton(x0, pgtol=1e-5, factr=1e6)
    # off again: the plain method (its iteration count varies a little from run to run: the
    # chain energies are summed over bead tiles with atomics, and thousands of line searches
    # amplify the last bits)
    assert np.all(it_off >= 5 * itr), (it_off, itr)


def test_lowest_eigenvalue_in_a_clustered_spectrum(evb):
    """lndetJ at instantons of the batched-instanton workload whose lowest eigenvalues lie a
    few per cent apart (three of its chains: N = 668, 499, 581 at their beta).  There inverse
    iteration from zero barely separates eta_1 from eta_2 and a Rayleigh-quotient shift can
    land on either, so eta_1 is held by an inertia bracket.  References: LAPACK's dsbevx
    (scipy.linalg.eig_banded, lowest eigenvalues only) on the band and the banded Cholesky's
    log-determinant; lndetJ to 1e-9, eta_1 (about 1e-6 of a band whose norm is about 1, so
    both factorisations fix it only to some 1e-15 absolute) to 1e-8, which still tells it
    from eta_2 (1.2e-4 above it at N = 668)."""
    sp, pes, op = evb
    ns = [668, 499, 581]
    betas = [1657.8751797669652, 1435.731658886481, 3046.54219889159]
    ring, betan = _ring(sp, pes, ns, betas)
    ring.set_preconditioner(0.05)
    xo, act, it, st = ring.instanton([_line(sp, n) for n in ns], pgtol=1e-5, factr=1e6)
    assert np.all(st == 0), (st, it)
    ring.set_x(xo)
    ld, nn, eta = ring.logdet(singlewell=False, skip_lowest=True)
    gaps = []
    for c, n in enumerate(ns):
        band = ring.UMhessian(c, singlewell=False)
        ev = sl.eig_banded(band, lower=True, eigvals_only=True, select="i", select_range=(0, 2))
        L = sl.cholesky_banded(band, lower=True)
        logdet = 2.0 * float(np.sum(np.log(L[0])))
        assert nn[c] == 0 and ev[0] > 0.0
        assert abs(eta[c] - ev[0]) <= 1e-8 * ev[0], (n, eta[c], ev)
        want = logdet - float(np.log(ev[0]))  # rpi_ser.f90:355-365: eta_1 skipped
        assert abs(ld[c] - want) <= 1e-9 * abs(want), (n, ld[c], want)
        gaps.append((ev[1] - ev[0]) / ev[0])
    assert min(gaps) < 0.05, gaps  # the case this test is for
    ring.close()
