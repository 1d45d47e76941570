"""Parity of the batched engine at the largest bead counts of the workloads, on the kernel
instances those workloads run, against the oracle.

The 27-dof EVB double well at N = 192 and 256 runs 12 and 16 MFMA units on 8 waves,
two units per wave, with folded transforms; the 6-water cluster (54 dof) at N = 128 runs
16 units on 8 waves with full transforms.  Each test first asserts that its families got
exactly those shapes by default, then holds every polymer to propagate_pimd_pile /
propagate_pimd_nm (verletmodule.f90:190-250, 372-416) for both thermostats, with injected
draws and with the engine's own Philox streams in blocks of 8 steps replayed from HIP
graphs.  Bars: states and TI sums 1e-10 normwise relative; stream cursors exact.
"""
import numpy as np
import pytest

import oracle as O
from helpers import oracle_pes, rand_chain, rel
from rptunnel import Batch, Pes, specs
from rptunnel.batch import ANDERSEN, LANGEVIN

pytestmark = pytest.mark.gpu

TAU, DT, GAMMA, BETA, NOUTPUT = 20.0, 1.0, 1e-3, 2000.0, 5


def _spec(name):
    return {"evb": specs.evb, "water6": lambda: specs.water_cluster(nmol=6)}[name]()


def _polymers(sp, ns, xis):
    pols = []
    for k, (n, xi) in enumerate(zip(ns, xis)):
        b = sp.well1 + xi * (sp.well2 - sp.well1)
        pols.append(dict(n=n, betan=BETA / (n + 1), xi=xi, a=sp.well1, b=b,
                         dbdl=sp.well2 - sp.well1, stream=300 + 7 * k))
    return pols


def _state(sp, n, seed):
    rng = np.random.default_rng(seed)
    x = rand_chain(sp, n, seed, scale=0.02)
    p = 0.003 * rng.standard_normal(x.shape) * np.sqrt(sp.mass.repeat(sp.ndim))[:, None]
    return x, p


# (potential, bead counts, expected (n, waves, units per wave, folded) per family)
CASES = [("evb", [192, 256, 192, 256], [(192, 8, 2, True), (256, 8, 2, True)]),
         ("water6", [128, 128], [(128, 8, 2, False)])]


def _make(sp, ns, thermostat, seed, xis=None):
    pes = Pes(sp)
    xis = xis or [0.15 + 0.7 * k / max(len(ns) - 1, 1) for k in range(len(ns))]
    pols = _polymers(sp, ns, xis)
    b = Batch(pes, sp.mass, pols, TAU, DT, GAMMA, thermostat=thermostat, noutput=NOUTPUT,
              seed=seed)
    states = []
    for k, pol in enumerate(pols):
        x, p = _state(sp, pol["n"], 500 + k)
        b.set_state(k, x, p)
        states.append((x, p))
    return pes, b, pols, states


def _nm(sp, pol):
    return O.NM(pol["n"], sp.ndim, sp.natom, sp.mass, pol["betan"], TAU, DT, GAMMA, False,
                pol["a"], pol["b"])


@pytest.mark.parametrize("name,ns,shapes", CASES)
@pytest.mark.parametrize("thermostat", [LANGEVIN, ANDERSEN])
def test_large_n_injected(name, ns, shapes, thermostat):
    sp = _spec(name)
    op = oracle_pes(sp)
    _, b, pols, st = _make(sp, ns, thermostat, seed=9)
    assert b.family_shapes() == shapes
    steps, imin = 24, 5
    rng = np.random.default_rng(17)
    if thermostat == LANGEVIN:
        noise = rng.standard_normal((len(pols), steps * max(ns) * sp.ndof))
        kicks = None
    else:
        kicks = rng.integers(1, 2 * NOUTPUT, size=(len(pols), steps + 2)).astype(np.int32)
        b.set_kicks(kicks)
        noise = rng.standard_normal((len(pols), kicks.shape[1] * max(ns) * sp.ndof))
    b.run(steps, imin=imin, noise=noise)
    sums, cnt, _ = b.ti_readout()
    for k, pol in enumerate(pols):
        nm = _nm(sp, pol)
        x, p = st[k]
        if thermostat == LANGEVIN:
            xo, po, dho, _ = O.propagate_pile(nm, op, O.Rng(normals=noise[k]), x, p, steps,
                                              pol["dbdl"], imin=imin)
        else:
            xo, po, dho = O.propagate_andersen(nm, op, O.Rng(normals=noise[k], kicks=kicks[k]),
                                               x, p, steps, pol["dbdl"], NOUTPUT, imin=imin)
        xg, pg = b.get_state(k)
        assert rel(xg, xo) <= 1e-10, (k, rel(xg, xo))
        assert rel(pg, po) <= 1e-10, (k, rel(pg, po))
        assert cnt[k] == steps - imin
        assert abs(sums[k] / (steps - imin) - dho) <= 1e-10 * abs(dho), (k, sums[k], dho)


@pytest.mark.parametrize("name,ns,shapes", CASES)
@pytest.mark.parametrize("thermostat", [LANGEVIN, ANDERSEN])
def test_large_n_philox_graph_blocks(name, ns, shapes, thermostat):
    """The benchmark's launch shape: Philox streams, 8 steps per launch replayed from
    captured graphs, three launches and a remainder block, stream cursors carried over."""
    sp = _spec(name)
    op = oracle_pes(sp)
    _, b, pols, st = _make(sp, ns, thermostat, seed=41)
    assert b.family_shapes() == shapes
    b.set_launch(steps_per_launch=8)
    b.run(24, imin=4)
    b.run(5)
    sums, cnt, _ = b.ti_readout()
    cur, kcur = b.cursors(), b.kick_cursors()
    for k, pol in enumerate(pols):
        nm = _nm(sp, pol)
        r = O.Rng(seed=41, stream=pol["stream"])
        x, p = st[k]
        if thermostat == LANGEVIN:
            xo, po, dho, _ = O.propagate_pile(nm, op, r, x, p, 29, pol["dbdl"], imin=4)
        else:
            xo, po, dho = O.propagate_andersen(nm, op, r, x, p, 29, pol["dbdl"], NOUTPUT, imin=4)
            assert kcur[k] == r.s.kcursor
        assert cur[k] == r.cursor
        xg, pg = b.get_state(k)
        assert rel(xg, xo) <= 1e-10, (k, rel(xg, xo))
        assert rel(pg, po) <= 1e-10, (k, rel(pg, po))
        assert cnt[k] == 25
        assert abs(sums[k] / 25 - dho) <= 1e-10 * abs(dho), (k, sums[k], dho)
