"""Parity of the batched engine (fused persistent step kernel) against the oracle.

Every polymer of a batch must follow propagate_pimd_pile / propagate_pimd_nm
(verletmodule.f90:190-250, 372-416) exactly: with injected draws, and with the
engine's own Philox streams (which the oracle replays).  Trajectories of 10-100 steps
are compared at 1e-10 normwise relative; TI sums at 1e-10 relative.
"""
import numpy as np
import pytest

import oracle as O
from helpers import oracle_pes, rand_chain, rel
from rptunnel import Batch, Chain, Pes, specs

pytestmark = pytest.mark.gpu

TAU, DT, GAMMA, BETA = 20.0, 1.0, 1e-3, 2000.0


def _polymers(sp, ns, xis=None):
    pols = []
    for k, n in enumerate(ns):
        xi = 0.5 if xis is None else xis[k]
        b = sp.well1 + xi * (sp.well2 - sp.well1)
        pols.append(dict(n=n, betan=BETA / (n + 1), xi=xi, a=sp.well1, b=b,
                         dbdl=sp.well2 - sp.well1, stream=100 + k))
    return pols


def _state(sp, n, seed):
    rng = np.random.default_rng(seed)
    x = rand_chain(sp, n, seed, scale=0.02)
    p = 0.3 * rng.standard_normal(x.shape) * np.sqrt(sp.mass.repeat(sp.ndim))[:, None] * 0.01
    return x, p


def _oracle_nm(sp, pol, cayley=False, gamma=GAMMA):
    return O.NM(pol["n"], sp.ndim, sp.natom, sp.mass, pol["betan"], TAU, DT, gamma, cayley,
                pol["a"], pol["b"])


@pytest.mark.parametrize("name,ns", [("evb", [32, 64, 96]), ("evb", [128, 40]),
                                     ("evb4", [64, 32]), ("evb10", [64, 32]),
                                     ("quartic", [64, 17]), ("coupled2d", [128, 32]),
                                     ("water6", [32]), ("water2", [64])])
def test_batch_pile_injected(name, ns):
    # evb4 / evb10: the kick's other padded shapes (12 dof in one 16-dof tile, 30 dof in two)
    sp = {"evb": specs.evb, "evb4": lambda: specs.evb(natom=4),
          "evb10": lambda: specs.evb(natom=10), "quartic": specs.quartic,
          "coupled2d": specs.coupled2d,
          "water6": lambda: specs.water_cluster(nmol=6),
          "water2": lambda: specs.water_cluster(nmol=2)}[name]()
    pes = Pes(sp)
    op = oracle_pes(sp)
    pols = _polymers(sp, ns)
    steps, imin = 24, 5
    b = Batch(pes, sp.mass, pols, TAU, DT, GAMMA, seed=9)
    cap = steps * max(ns) * sp.ndof
    noise = np.random.default_rng(1).standard_normal((len(pols), cap))
    states = []
    for k, pol in enumerate(pols):
        x, p = _state(sp, pol["n"], 10 + k)
        b.set_state(k, x, p)
        states.append((x, p))
    b.run(steps, imin=imin, noise=noise)
    sums, cnt, _ = b.ti_readout()
    for k, pol in enumerate(pols):
        nm = _oracle_nm(sp, pol)
        x, p = states[k]
        xo, po, dho, _ = O.propagate_pile(nm, op, O.Rng(normals=noise[k]), x, p, steps,
                                          pol["dbdl"], imin=imin)
        xg, pg = b.get_state(k)
        assert rel(xg, xo) <= 1e-10, (k, rel(xg, xo))
        assert rel(pg, po) <= 1e-10, (k, rel(pg, po))
        assert cnt[k] == steps - imin
        assert abs(sums[k] / (steps - imin) - dho) <= 1e-10 * abs(dho)


def test_batch_pile_philox_and_blocks():
    """Philox streams, several launches (steps per launch < steps), Cayley off/on."""
    sp = specs.evb()
    pes = Pes(sp)
    op = oracle_pes(sp)
    for cayley in (False, True):
        pols = _polymers(sp, [64, 32, 64])
        b = Batch(pes, sp.mass, pols, TAU, DT, GAMMA, cayley=cayley, seed=77)
        b.set_launch(steps_per_launch=7)
        st = []
        for k, pol in enumerate(pols):
            x, p = _state(sp, pol["n"], 20 + k)
            b.set_state(k, x, p)
            st.append((x, p))
        b.run(30)
        cur = b.cursors()
        for k, pol in enumerate(pols):
            nm = _oracle_nm(sp, pol, cayley=cayley)
            r = O.Rng(seed=77, stream=pol["stream"])
            xo, po, dho, _ = O.propagate_pile(nm, op, r, st[k][0], st[k][1], 30, pol["dbdl"])
            xg, pg = b.get_state(k)
            assert cur[k] == r.cursor
            assert rel(xg, xo) <= 1e-10 and rel(pg, po) <= 1e-10


@pytest.mark.parametrize("thermostat", [2, 1])
def test_graph_replays_match_direct_launches(thermostat):
    """Step blocks replayed from captured HIP graphs give bit for bit the states, TI sums,
    counts and stream positions of direct launches, over several runs of different lengths
    (a remainder block, a changed imin) with host reads between the replays."""
    sp = specs.evb()
    pes = Pes(sp)
    pols = _polymers(sp, [64, 32, 96, 64], xis=[0.2, 0.5, 0.7, 0.9])
    bs = []
    for graphs in (True, False):
        b = Batch(pes, sp.mass, pols, TAU, DT, GAMMA, thermostat=thermostat, noutput=5, seed=21)
        b.set_graphs(graphs)
        b.set_launch(steps_per_launch=8)
        for k, pol in enumerate(pols):
            x, p = _state(sp, pol["n"], 40 + k)
            b.set_state(k, x, p)
        bs.append(b)
    done = 0
    for nsteps, imin in ((17, 3), (16, 0), (9, 0), (8, 0)):
        for b in bs:
            b.run(nsteps, imin=imin)
        done += nsteps
        for k in range(len(pols)):
            xg, pg = bs[0].get_state(k)
            xe, pe = bs[1].get_state(k)
            assert np.array_equal(xg, xe) and np.array_equal(pg, pe), (nsteps, k)
        sg, cg, _ = bs[0].ti_readout()
        se, ce, _ = bs[1].ti_readout()
        assert np.array_equal(sg, se) and np.array_equal(cg, ce)
        assert np.all(cg == done - 3)  # the step counter carries over between replays
        assert np.array_equal(bs[0].cursors(), bs[1].cursors())
        assert np.array_equal(bs[0].kick_cursors(), bs[1].kick_cursors())


@pytest.mark.parametrize("thermostat", [2, 1])
def test_wave_counts_agree(thermostat):
    """Every workgroup shape gives bit for bit the same states, TI sums and stream positions:
    2 waves (families of at most 2 MFMA units: N = 24, 32), 4 waves and 8 (families of at
    most 16 units, here all): idle unit slots at N = 24, 32, 33 and 64, and two units per
    wave at N = 128 with 4 waves."""
    sp = specs.evb()
    pes = Pes(sp)
    pols = _polymers(sp, [32, 24, 64, 128, 33], xis=[0.2, 0.5, 0.7, 0.9, 0.4])
    b = Batch(pes, sp.mass, pols, TAU, DT, GAMMA, thermostat=thermostat, noutput=5, seed=13)
    default = {24: 2, 32: 2, 33: 4, 64: 4, 128: 8}  # units: 2, 2, 4, 4, 8
    assert all(s[1] == default[n] for (n, _), s in zip(b.families(), b.family_shapes()))
    b.close()
    res = []
    for waves in (4, 8, 2):
        b = Batch(pes, sp.mass, pols, TAU, DT, GAMMA, thermostat=thermostat, noutput=5, seed=13)
        b.set_launch(steps_per_launch=8, waves=waves)
        want = {24: 2, 32: 2, 33: 4, 64: 4, 128: 4} if waves == 2 else None
        for (n, _), shape in zip(b.families(), b.family_shapes()):
            assert shape[1] == (want[n] if want else waves), (n, shape)
        for k, pol in enumerate(pols):
            x, p = _state(sp, pol["n"], 60 + k)
            b.set_state(k, x, p)
        b.run(20, imin=3)
        sums, cnt, _ = b.ti_readout()
        res.append(([b.get_state(k) for k in range(len(pols))], sums, cnt, b.cursors(),
                    b.kick_cursors()))
    for r in res[1:]:
        for k in range(len(pols)):
            assert np.array_equal(r[0][k][0], res[0][0][k][0]), k
            assert np.array_equal(r[0][k][1], res[0][0][k][1]), k
        for i in (1, 2, 3, 4):
            assert np.array_equal(r[i], res[0][i])
    assert np.all(res[0][2] == 17)


def test_batch_init_path_and_reinit():
    sp = specs.evb()
    pes = Pes(sp)
    op = oracle_pes(sp)
    lampath, path = specs.linear_path(sp.well1, sp.well2, npath=7)
    path = path + 0.02 * np.sin(np.arange(7))[None, :]
    opath = O.Path(lampath, path)
    pols = _polymers(sp, [32, 64], xis=[0.6, 0.3])
    b = Batch(pes, sp.mass, pols, TAU, DT, GAMMA, dhdrlimit=300.0, seed=5)
    b.set_path(lampath, path)
    b.init_path()
    # init_path from the Philox streams
    inits = []
    for k, pol in enumerate(pols):
        nm = _oracle_nm(sp, pol)
        r = O.Rng(seed=5, stream=pol["stream"])
        xo, po = O.init_path(nm, opath, r, pol["xi"])
        xg, pg = b.get_state(k)
        assert rel(xg, xo) <= 1e-12 and rel(pg, po) <= 1e-12
        inits.append((xo, po, r.cursor))
    # dHdrlimit re-initialisations mid-run
    b.run(60)
    sums, cnt, nre = b.ti_readout()
    for k, pol in enumerate(pols):
        nm = _oracle_nm(sp, pol)
        r = O.Rng(seed=5, stream=pol["stream"])
        r.s.cursor = inits[k][2]
        xo, po, dho, nreo = O.propagate_pile(nm, op, r, inits[k][0], inits[k][1], 60, pol["dbdl"],
                                             dhdrlimit=300.0, xi=pol["xi"], path=opath)
        assert nre[k] == nreo
        xg, pg = b.get_state(k)
        assert rel(xg, xo) <= 1e-10
        assert abs(sums[k] / 60 - dho) <= 1e-10 * max(abs(dho), 1e-300)


def test_batch_andersen_injected():
    sp = specs.quartic()
    pes = Pes(sp)
    op = oracle_pes(sp)
    pols = _polymers(sp, [64, 32])
    b = Batch(pes, sp.mass, pols, TAU, DT, thermostat=1, noutput=5, seed=3)
    kicks = np.array([[3, 5, 2, 7, 4, 6, 3, 5, 2, 8, 4, 4, 9, 3, 5],
                      [2, 2, 3, 9, 1, 4, 6, 2, 5, 3, 3, 7, 2, 2, 4]], dtype=np.int32)
    b.set_kicks(kicks)
    cap = kicks.shape[1] * 64 * sp.ndof
    noise = np.random.default_rng(4).standard_normal((2, cap))
    st = []
    for k, pol in enumerate(pols):
        x, p = _state(sp, pol["n"], 30 + k)
        b.set_state(k, x, p)
        st.append((x, p))
    b.run(40, imin=4, noise=noise)
    sums, cnt, _ = b.ti_readout()
    for k, pol in enumerate(pols):
        nm = _oracle_nm(sp, pol)
        xo, po, dho = O.propagate_andersen(nm, op, O.Rng(normals=noise[k], kicks=kicks[k]),
                                           st[k][0], st[k][1], 40, pol["dbdl"], 5, imin=4)
        xg, pg = b.get_state(k)
        assert rel(xg, xo) <= 1e-10 and rel(pg, po) <= 1e-10
        assert abs(sums[k] / 36 - dho) <= 1e-10 * abs(dho)


def test_batch_equals_plain_path():
    """The drop-in plain path and the batched engine run the same trajectory."""
    sp = specs.evb()
    pes = Pes(sp)
    pol = _polymers(sp, [64])[0]
    x, p = _state(sp, 64, 40)
    b = Batch(pes, sp.mass, [pol], TAU, DT, GAMMA, seed=123)
    b.set_state(0, x, p)
    b.run(50, imin=10)
    sums, _, _ = b.ti_readout()
    ch = Chain(pes, 64, sp.mass, pol["betan"], TAU, DT, GAMMA, pol["a"], pol["b"])
    ch.set_state(x, p)
    ch.set_rng(123, pol["stream"])
    dh, _ = ch.propagate_pimd_pile(50, pol["dbdl"], imin=10)
    xg, _ = b.get_state(0)
    xc, _ = ch.get_state()
    assert rel(xg, xc) <= 1e-10
    assert abs(sums[0] / 40 - dh) <= 1e-10 * abs(dh)


def test_autotune_keeps_results_and_state():
    """rp_batch_autotune times launch shapes on the live batch and restores its state: the
    state, TI sums and stream positions afterwards are those before it, and runs with the
    configuration it picks give bit for bit the default configuration's states, counts and
    stream positions (TI sums to rounding: they are grouped per launch)."""
    sp = specs.evb()
    pes = Pes(sp)
    pols = _polymers(sp, [32, 128, 32, 96], xis=[0.2, 0.5, 0.7, 0.9])
    bs = []
    for tune in (False, True):
        b = Batch(pes, sp.mass, pols, TAU, DT, GAMMA, seed=31)
        for k, pol in enumerate(pols):
            x, p = _state(sp, pol["n"], 80 + k)
            b.set_state(k, x, p)
        b.run(5, imin=1)
        if tune:
            before = [b.get_state(k) for k in range(len(pols))], b.ti_readout(), b.cursors()
            steps, waves = b.autotune(trial_steps=8)
            assert steps >= 1 and waves in (2, 4, 8)
            after = [b.get_state(k) for k in range(len(pols))], b.ti_readout(), b.cursors()
            for (x0, p0), (x1, p1) in zip(before[0], after[0]):
                assert np.array_equal(x0, x1) and np.array_equal(p0, p1)
            assert all(np.array_equal(u, v) for u, v in zip(before[1], after[1]))
            assert np.array_equal(before[2], after[2])
        b.run(19)
        bs.append(b)
    for k in range(len(pols)):
        xa, pa = bs[0].get_state(k)
        xb, pb = bs[1].get_state(k)
        assert np.array_equal(xa, xb) and np.array_equal(pa, pb), k
    (s0, c0, r0), (s1, c1, r1) = bs[0].ti_readout(), bs[1].ti_readout()
    assert np.array_equal(c0, c1) and np.array_equal(r0, r1)
    # a launch adds its steps' TI terms to the running sum as one partial sum, so another
    # number of steps per launch groups the same terms differently (rounding only)
    assert np.all(np.abs(s0 - s1) <= 1e-13 * np.abs(s0))
    assert np.array_equal(bs[0].cursors(), bs[1].cursors())
