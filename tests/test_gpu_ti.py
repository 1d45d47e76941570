"""Thermodynamic integration end to end on the batched engine, against closed forms.

For V = k (x - c)^2 / 2 the open chain with fixed ends a and b(xi) is Gaussian, and its mean
solves (m/betan^2 L + k) xbar = m/betan^2 (a e_1 + b(xi) e_N) + k c.  The reference's TI
integrand (verletmodule.f90:397-410 with pimd_par.f90:379) is <sum m (-x_N) db/dxi> / betan^2,
so the quadrature Delta A = sum_i w_i integrand(xi_i) (pimd_par.f90:397-420) is known
exactly.  run_ti drives the batch the way the reference's task loop does (node-major
tasks, init_path at each node, imin equilibration steps); its Delta A must agree within
four of its own standard errors, for both thermostats."""
import math

import numpy as np
import pytest

from rptunnel import Batch, Pes, ti
from rptunnel.batch import ANDERSEN, LANGEVIN
from rptunnel.specs import PES_HARMONIC, PesSpec

pytestmark = pytest.mark.gpu


def _lap(n):
    return 2.0 * np.eye(n) - np.eye(n, k=1) - np.eye(n, k=-1)


@pytest.mark.parametrize("thermostat", [LANGEVIN, ANDERSEN])
def test_ti_harmonic_delta_a(thermostat):
    m, omega, c, a, b1 = 1836.0, 0.01, 0.1, -0.3, 0.5
    k = m * omega ** 2
    sp = PesSpec("harmonic-ti", PES_HARMONIC, 1, 1, np.array([k, c]), None, np.array([m]),
                 np.array([a]), np.array([b1]))
    pes = Pes(sp)
    n, beta, nint, nrep, nsteps, imin = 8, 1000.0, 5, 64, 20000, 2000
    betan = beta / (n + 1)

    def factory(pols):
        return Batch(pes, sp.mass, pols, 20.0, 1.0, gamma=0.05, thermostat=thermostat,
                     noutput=40, seed=99)

    res = ti.run_ti(factory, sp, n, beta, nint, nrep, nsteps, imin)
    # closed form at the same Gauss-Legendre nodes
    A = (m / betan ** 2) * _lap(n) + k * np.eye(n)
    want = 0.0
    for xi, w in zip(res["nodes"], res["weights"]):
        b = a + xi * (b1 - a)
        rhs = np.full(n, k * c)
        rhs[0] += m * a / betan ** 2
        rhs[-1] += m * b / betan ** 2
        xbar = np.linalg.solve(A, rhs)
        want += w * (-m * xbar[-1] * (b1 - a)) / betan ** 2
    err = math.sqrt(res["sigmaA"] / nrep)
    assert abs(res["deltaA"] - want) <= 4.0 * err, (res["deltaA"], want, err)
    assert err <= 0.05 * abs(want)  # the comparison has power
    assert res["q_ratio"] == pytest.approx(math.exp(-res["deltaA"] * betan), rel=1e-14)
