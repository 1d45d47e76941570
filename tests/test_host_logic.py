"""Host-side logic of rptunnel that runs without a GPU: the TI driver arithmetic
(pimd_par.f90:243-424), the splitting formulas (rpi_ser.f90:368-381,
multiwell_waterdimer.py:6-16), the path spline used for b(xi) and db/dxi
(instantonmod.f90:397-561) and the synthetic potentials' invariants."""
import math

import numpy as np
import pytest
import scipy.interpolate

import oracle as O
from helpers import oracle_pes, rel, spec_by_name
from rptunnel import specs, ti


def test_gauleg_matches_oracle_and_numpy():
    for n in (1, 3, 11, 16):
        x, w = ti.gauleg(0.0, 1.0, n)
        xo, wo = O.gauleg(0.0, 1.0, n)
        assert np.array_equal(x, xo) or np.abs(x - xo).max() <= 1e-15
        assert np.abs(w - wo).max() <= 1e-15
        xr, wr = np.polynomial.legendre.leggauss(n)
        assert np.abs(x - (xr + 1) / 2).max() <= 1e-14


def test_task_table_and_skip():
    tab = ti.task_table(3, 4)
    assert tab[0] == (0, 0) and tab[4] == (1, 0) and len(tab) == 12  # lambda-major
    assert ti.skip_task(np.array([0.0, 5e-11]))
    assert not ti.skip_task(np.array([0.0, 2e-10]))


def test_quadrature_matches_oracle_restatement():
    rng = np.random.default_rng(1)
    _, w = ti.gauleg(0.0, 1.0, 5)
    I = rng.standard_normal(5 * 7) * 0.01 + 0.02
    I[7:14] = np.nan  # a node whose replicas all failed is dropped from Delta A
    a = ti.quadrature(I, w, 7, 30.0)
    b = O.ti_quadrature(I, w, 7, 30.0)
    assert a["deltaA"] == pytest.approx(b["deltaA"], rel=1e-14)
    assert a["q_ratio"] == pytest.approx(b["q_ratio"], rel=1e-14)
    # the population variance of a node with identical replicas is zero
    J = np.repeat([0.1, 0.2, 0.3, 0.4, 0.5], 7)
    c = ti.quadrature(J, w, 7, 30.0)
    assert c["sigmaA"] == pytest.approx(0.0, abs=1e-16)  # E[x^2] - E[x]^2 rounding
    assert c["deltaA"] == pytest.approx(np.dot(w, [0.1, 0.2, 0.3, 0.4, 0.5]), rel=1e-14)
    assert c["q_ratio"] == pytest.approx(math.exp(-c["deltaA"] * 30.0), rel=1e-14)


def test_two_beta_splitting_recovers_a_two_level_system():
    """For a symmetric two-level system q/q0 = tanh(beta Delta / 2) is the ratio of the
    antisymmetric to the symmetric partition function difference; the formula inverts it."""
    delta = 3.7e-5
    b1, b2 = 20000.0, 35000.0
    I1, I2 = math.tanh(b1 * delta / 2 + 0.01), math.tanh(b2 * delta / 2 + 0.01)
    assert ti.splitting_two_beta(I1, I2, b1, b2) == pytest.approx(delta, rel=1e-10)
    assert O.splitting_two_beta(I1, I2, b1, b2) == pytest.approx(delta, rel=1e-10)
    assert ti.splitting_single_beta(math.tanh(b1 * delta / 2), b1) == pytest.approx(delta, rel=1e-12)


def test_rpi_splitting_matches_oracle():
    a = ti.rpi_splitting(40.0, 9.5, -3.25, -2.0)
    b = O.rpi_splitting(40.0, 9.5, -3.25, -2.0)
    assert np.allclose(a, b, rtol=1e-15, atol=0)


def test_path_nodes_against_scipy():
    rng = np.random.default_rng(2)
    path = rng.standard_normal((4, 7))
    lam = specs.arc_lampath(path)
    xi, _ = ti.gauleg(0.0, 1.0, 11)
    ends, grads = ti._spline_nodes(lam, path, xi)
    cs = scipy.interpolate.CubicSpline(lam, path, axis=1, bc_type="natural")
    assert rel(ends, cs(xi).T) <= 1e-12
    assert rel(grads, cs(xi, 1).T) <= 1e-11
    op = O.Path(lam, path)
    for k in (0, 5, 10):
        assert rel(ends[k], op.splint(xi[k])) <= 1e-14
        assert rel(grads[k], op.splin_grad(xi[k])) <= 1e-14


def test_linear_path_is_arc_length_normalised():
    sp = specs.coupled2d()
    lam, path = specs.linear_path(sp.well1, sp.well2, npath=5)
    assert lam[0] == 0.0 and lam[-1] == 1.0 and np.allclose(np.diff(lam), 0.25)
    assert np.allclose(path[:, 0], sp.well1) and np.allclose(path[:, -1], sp.well2)


@pytest.mark.parametrize("name", ["quartic", "coupled2d", "evb"])
def test_synthetic_double_wells_are_symmetric_minima(name):
    sp = spec_by_name(name)
    op = oracle_pes(sp)
    v1, v2 = op.V(sp.well1), op.V(sp.well2)
    assert v1 == pytest.approx(v2, abs=1e-15)
    for w in (sp.well1, sp.well2):
        g = op.Vprime(w)
        if name == "evb":
            # the EVB ends are the diabatic minima x_A and Pi x_A; there the coupling
            # adds a rank-one softening, so they are near-stationary but not minima
            assert np.abs(g).max() <= 1e-3
            continue
        assert np.abs(g).max() <= 1e-12
        assert np.linalg.eigvalsh(op.Vdoubleprime(w)).min() > 0
    if name == "evb":
        gp = op.Vprime(sp.well2)
        perm = np.repeat(sp.ip * 3, 3) + np.tile(np.arange(3), sp.natom)
        assert rel(gp, op.Vprime(sp.well1)[perm]) <= 1e-12  # V(Pi x) = V(x)
    mid = 0.5 * (sp.well1 + sp.well2)
    assert op.V(mid) > v1  # a barrier between them


def test_evb_spec_is_reproducible_and_orthogonal():
    a, b = specs.evb(), specs.evb()
    assert np.array_equal(a.dp, b.dp) and np.array_equal(a.ip, b.ip)
    Q = a.extra["Q"]
    assert np.abs(Q.T @ Q - np.eye(27)).max() <= 1e-13
    assert a.mass[0] == a.mass[1]  # the permuted atoms share a mass
