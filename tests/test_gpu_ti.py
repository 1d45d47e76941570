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


def test_quartic_ti_pipeline_matches_oracle():
    """The 1D quartic double well (V0 = 0.005, a = 0.7, m = 1836), N = 64, beta = 2000,
    11 Gauss-Legendre nodes x 2 replicas, driven the way the reference's task loop does
    (pimd_par.f90:321-381, 397-424): node-major tasks, skipped b = 0 tasks, a linear
    initial path placed by init_path (verletmodule.f90:32-119), imin equilibration steps,
    the TI readout and the quadrature.  Every polymer replays on the oracle with the same Philox stream.
    Rounding differences grow by about 10x per 100 steps in this well (2e-11 after 200
    steps, 5e-7 after 1000), so 300 steps are compared per task at 1e-8 relative."""
    import oracle as O
    from helpers import oracle_pes
    from rptunnel import specs

    sp = specs.quartic(V0=0.005, a=0.7, mass=1836.0)
    pes, op = Pes(sp), oracle_pes(sp)
    n, beta, nint, nrep, nsteps, imin, seed = 64, 2000.0, 11, 2, 300, 50, 11
    tau, dt, gamma = 20.0, 1.0, 1e-3
    betan = beta / (n + 1)
    lampath, path = specs.linear_path(sp.well1, sp.well2, npath=2)
    xi, w = ti.gauleg(0.0, 1.0, nint)
    pols, keep = [], []
    for k, (i, _) in enumerate(ti.task_table(nint, nrep)):
        b = sp.well1 + xi[i] * (sp.well2 - sp.well1)
        if ti.skip_task(b):
            continue
        pols.append(dict(n=n, betan=betan, xi=float(xi[i]), a=sp.well1, b=b,
                         dbdl=sp.well2 - sp.well1, stream=k))
        keep.append(k)
    # the middle node of the symmetric well sits at b = 0: its replicas are skipped, integrand 0
    assert len(pols) == (nint - 1) * nrep
    bt = Batch(pes, sp.mass, pols, tau, dt, gamma, seed=seed)
    bt.set_path(lampath, path)
    bt.init_path()
    bt.run(nsteps, imin=imin)
    sums, cnt, _ = bt.ti_readout()
    assert np.all(cnt == nsteps - imin)
    got = sums / (nsteps - imin) / betan ** 2
    opath = O.Path(lampath, path)
    want = np.empty(len(pols))
    for q, pol in enumerate(pols):
        nm = O.NM(n, 1, 1, sp.mass, betan, tau, dt, gamma, False, pol["a"], pol["b"])
        r = O.Rng(seed=seed, stream=pol["stream"])
        xo, po = O.init_path(nm, opath, r, pol["xi"])
        _, _, dho, _ = O.propagate_pile(nm, op, r, xo, po, nsteps, pol["dbdl"], imin=imin)
        want[q] = dho / betan ** 2
    assert np.max(np.abs(got - want) / np.abs(want)) <= 1e-8
    ig, iw = np.zeros(nint * nrep), np.zeros(nint * nrep)
    ig[keep], iw[keep] = got, want
    rg, rw = ti.quadrature(ig, w, nrep, betan), ti.quadrature(iw, w, nrep, betan)
    assert rg["deltaA"] == pytest.approx(rw["deltaA"], rel=1e-8)
    assert rg["q_ratio"] == pytest.approx(rw["q_ratio"], rel=1e-8)
