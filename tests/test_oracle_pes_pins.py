"""Pins of the oracle's potentials by values the mathematics fixes (not by their own formulas).

None of the workload potentials exists in the reference (each is a new mcmod_mass plugin
against the interface of mcmod_waterdimer.f90:12-103), so their energies are held to
special points and limits, chosen so that a plausible slip -- a wrong constant, a
transposed matrix, a dropped or mis-signed term, a wrong charge -- fails one of them:

* quartic (mcmod_1d.f90:20's form): V(0) = V0 per dof, V(+-a) = 0, V(a/2) = 9 V0 / 16.
* coupled 2D: V(0, 0) = V0, zero at the minima (+-a, -c a^2), a pure parabola in y at
  x = 0, and the quartic alone along the valley y = -c x^2.
* EVB: with the identity permutation both diabats coincide, so V = VA - kappa exactly,
  and along an eigenvector of well A, x = xA + t Q e_k, VA = lam_k t^2 / 2 + c_k t^3.
  That gives VA(z) at any z, and with the atom swap V(x) must be the lower eigenvalue
  (numpy's eigvalsh) of the 2 x 2 EVB matrix [[VA(x), kappa], [kappa, VA(Pi x)]].
* water: a monomer at (re, re, theta0) has V = 0; one stretched bond gives the Morse
  value De (1 - exp(-alpha d))^2; a bent angle gives kbend phi^2 / 2.  For two molecules
  far apart the interaction times R^3 tends to the point-dipole energy
  mu_A.mu_B - 3 (mu_A.R^)(mu_B.R^) (which holds only if qO = -2 qH makes each
  molecule neutral); at contact distances the interaction equals the pair sum of the
  point charges plus the O-O Lennard-Jones term.
* dV/dqH and dV/dDe (the adjoint parameter sensitivities): V is a quadratic in qH and
  linear in De, so central differences in the parameter are exact up to rounding.
"""
import math

import numpy as np
import pytest

import oracle as O
from rptunnel import specs


def _pes(sp, dp=None, ip=None):
    return O.Pes(sp.kind, sp.ndim, sp.natom, sp.dp if dp is None else dp,
                 sp.ip if ip is None else ip)


# ----------------------------------------------------------------------------- quartic
def test_quartic_special_points():
    V0, a = 0.005, 0.7
    for ndim in (1, 3):
        op = _pes(specs.quartic(V0=V0, a=a, ndim=ndim))
        assert op.V(np.zeros(ndim)) == pytest.approx(ndim * V0, rel=1e-15)
        assert abs(op.V(np.full(ndim, a))) <= 1e-18
        assert abs(op.V(np.full(ndim, -a))) <= 1e-18
        assert op.V(np.full(ndim, a / 2)) == pytest.approx(ndim * 9.0 * V0 / 16.0, rel=1e-14)
    # V0 offset: the module variable V0 of mcmod_waterdimer.f90:3 is subtracted
    sp = specs.quartic(V0=V0, a=a)
    op = O.Pes(sp.kind, sp.ndim, sp.natom, sp.dp, None, v0=0.001)
    assert op.V(np.zeros(1)) == pytest.approx(V0 - 0.001, rel=1e-14)


# ----------------------------------------------------------------------------- coupled 2D
def test_coupled2d_special_points():
    V0, a, w, c, m = 0.005, 0.7, 0.01, 0.3, 1836.0
    op = _pes(specs.coupled2d(V0=V0, a=a, w=w, c=c, m=m))
    assert op.V([0.0, 0.0]) == pytest.approx(V0, rel=1e-15)
    for s in (1.0, -1.0):
        assert abs(op.V([s * a, -c * a * a])) <= 1e-18
    for y in (0.1, -0.37):
        assert op.V([0.0, y]) == pytest.approx(V0 + 0.5 * m * w * w * y * y, rel=1e-14)
    for x in (0.2, -0.9):
        assert op.V([x, -c * x * x]) == pytest.approx(V0 * (x * x / (a * a) - 1.0) ** 2, rel=1e-13)


# ----------------------------------------------------------------------------- EVB
def _evb_identity(sp):
    return _pes(sp, ip=np.arange(sp.natom, dtype=np.int32))


def test_evb_single_diabat_along_eigenvectors():
    """Identity permutation: V = VA - kappa, and VA along Q e_k is lam_k t^2/2 + c_k t^3."""
    sp = specs.evb()
    Q, lam, cub = sp.extra["Q"], sp.extra["lam"], sp.extra["cub"]
    kappa = sp.dp[0]
    op = _evb_identity(sp)
    assert op.V(sp.well1) == pytest.approx(-kappa, rel=1e-14)
    for k in (0, 5, 26):
        for t in (0.3, -0.8):
            want = 0.5 * lam[k] * t * t + cub[k] * t ** 3 - kappa
            assert op.V(sp.well1 + t * Q[:, k]) == pytest.approx(want, rel=1e-12, abs=1e-17)
    # two directions at once: the cross terms vanish for an orthogonal Q
    t0, t1 = 0.4, -0.25
    x = sp.well1 + t0 * Q[:, 3] + t1 * Q[:, 20]
    want = 0.5 * (lam[3] * t0 ** 2 + lam[20] * t1 ** 2) + cub[3] * t0 ** 3 + cub[20] * t1 ** 3 - kappa
    assert op.V(x) == pytest.approx(want, rel=1e-12)


def test_evb_is_lower_eigenvalue_of_two_state_matrix():
    sp = specs.evb()
    kappa = sp.dp[0]
    one = _evb_identity(sp)
    op = _pes(sp)
    perm = np.repeat(sp.ip * 3, 3) + np.tile(np.arange(3), sp.natom)
    rng = np.random.default_rng(3)
    pts = [sp.well1, sp.well2, 0.5 * (sp.well1 + sp.well2)]
    pts += [sp.well1 + t * (sp.well2 - sp.well1) + 0.05 * rng.standard_normal(sp.ndof)
            for t in (0.2, 0.45, 0.8)]
    for x in pts:
        va = one.V(x) + kappa
        vb = one.V(x[perm]) + kappa  # VB(x) = VA(Pi x)
        lo = np.linalg.eigvalsh(np.array([[va, kappa], [kappa, vb]]))[0]
        assert op.V(x) == pytest.approx(lo, rel=1e-12, abs=1e-16)
    # on the swap-symmetric midpoint the diabats are degenerate: V = VA - kappa
    mid = 0.5 * (sp.well1 + sp.well2)
    assert op.V(mid) == pytest.approx(one.V(mid), rel=1e-13)


# ----------------------------------------------------------------------------- water
def _monomer(re, theta, r1=None, r2=None):
    r1 = re if r1 is None else r1
    r2 = re if r2 is None else r2
    O_ = np.zeros(3)
    H1 = np.array([r1, 0.0, 0.0])
    H2 = r2 * np.array([math.cos(theta), math.sin(theta), 0.0])
    return np.concatenate([O_, H1, H2])


def _water(nmol, **kw):
    sp = specs.water_cluster(nmol=nmol, **kw)
    return sp, _pes(sp)


def test_water_monomer_intramolecular_terms():
    sp, op = _water(1)
    De, al, re, kb, th0 = sp.dp[:5]
    assert abs(op.V(_monomer(re, th0))) <= 1e-15
    assert np.abs(op.Vprime(_monomer(re, th0))).max() <= 1e-14
    for d in (0.2, -0.15):
        want = De * (1.0 - math.exp(-al * d)) ** 2
        assert op.V(_monomer(re, th0, r1=re + d)) == pytest.approx(want, rel=1e-12)
        assert op.V(_monomer(re, th0, r2=re + d)) == pytest.approx(want, rel=1e-12)
    for phi in (0.1, -0.3):
        assert op.V(_monomer(re, th0 + phi)) == pytest.approx(0.5 * kb * phi * phi, rel=1e-10)
    # a rigid motion leaves V unchanged
    rng = np.random.default_rng(1)
    A, _ = np.linalg.qr(rng.standard_normal((3, 3)))
    x = _monomer(re + 0.1, th0 - 0.2).reshape(3, 3)
    y = (x @ A.T + np.array([1.0, -2.0, 0.5])).reshape(-1)
    assert op.V(y) == pytest.approx(op.V(x.reshape(-1)), rel=1e-13)


def _dimer(sep, rot):
    sp, op = _water(2)
    re, th0 = sp.dp[2], sp.dp[4]
    A = _monomer(re + 0.05, th0 + 0.1).reshape(3, 3)
    B = (_monomer(re - 0.03, th0 - 0.05).reshape(3, 3) @ rot.T) + sep
    return sp, op, A, B


def _rot(seed):
    Qm, R = np.linalg.qr(np.random.default_rng(seed).standard_normal((3, 3)))
    return Qm * np.sign(np.diag(R))


def _interaction(op, mono, A, B):
    return op.V(np.concatenate([A.reshape(-1), B.reshape(-1)])) - mono.V(A.reshape(-1)) - mono.V(B.reshape(-1))


def test_water_far_dimer_is_point_dipole():
    _, mono = _water(1)
    rot = _rot(5)
    R_hat = np.array([0.6, 0.0, 0.8])
    vals = []
    for R in (400.0, 800.0):
        sp, op, A, B = _dimer(R * R_hat, rot)
        qH = sp.dp[7]
        q = np.array([-2.0 * qH, qH, qH])
        muA = q @ A
        muB = q @ (B - R * R_hat)
        want = (muA @ muB - 3.0 * (muA @ R_hat) * (muB @ R_hat)) / R ** 3
        vals.append((_interaction(op, mono, A, B), want))
    # relative quadrupole corrections fall as 1/R: halve from 400 to 800 bohr
    e1 = abs(vals[0][0] / vals[0][1] - 1.0)
    e2 = abs(vals[1][0] / vals[1][1] - 1.0)
    assert e1 <= 0.01 and e2 <= 0.55 * e1


def test_water_contact_dimer_is_charge_pairs_plus_lj():
    _, mono = _water(1)
    rot = _rot(9)
    sp, op, A, B = _dimer(np.array([5.4, 0.7, -0.4]), rot)
    eps, sig, qH = sp.dp[5], sp.dp[6], sp.dp[7]
    q = np.array([-2.0 * qH, qH, qH])
    coul = sum(q[i] * q[j] / np.linalg.norm(A[i] - B[j]) for i in range(3) for j in range(3))
    roo = np.linalg.norm(A[0] - B[0])
    lj = 4.0 * eps * ((sig / roo) ** 12 - (sig / roo) ** 6)
    assert _interaction(op, mono, A, B) == pytest.approx(coul + lj, rel=1e-11)


@pytest.mark.parametrize("nmol", [2, 6])
def test_water_parameter_gradient_by_central_differences(nmol):
    sp = specs.water_cluster(nmol=nmol)
    rng = np.random.default_rng(nmol)
    x = sp.well1 + 0.03 * rng.standard_normal(sp.ndof)
    op = _pes(sp)
    pg = op.param_grad(x)
    assert pg.shape == (2,)
    for k, idx, h in ((0, 7, 1e-3), (1, 0, 1e-3)):  # (dV/dqH, dp[7]), (dV/dDe, dp[0])
        dpp, dpm = sp.dp.copy(), sp.dp.copy()
        dpp[idx] += h
        dpm[idx] -= h
        fd = (_pes(sp, dp=dpp).V(x) - _pes(sp, dp=dpm).V(x)) / (2.0 * h)
        assert pg[k] == pytest.approx(fd, rel=1e-9), (k, pg[k], fd)
    # V is quadratic in qH with no linear part: V(qH) - V(0) = qH dV/dqH / 2
    d0 = sp.dp.copy()
    d0[7] = 0.0
    assert op.V(x) - _pes(sp, dp=d0).V(x) == pytest.approx(0.5 * sp.dp[7] * pg[0], rel=1e-12)
