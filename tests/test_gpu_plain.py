"""Parity of the plain per-trajectory GPU path against the oracle (FP64).

Tolerance: 1e-12 normwise relative for energies, forces, normal-mode transforms and
single Verlet steps (the bar the reference comparison is set at); multi-step
trajectories 1e-10 (rounding differences grow along a trajectory).
"""
import numpy as np
import pytest

import oracle as O
from helpers import oracle_pes, rand_chain, rel
from rptunnel import Chain, Pes, specs

pytestmark = pytest.mark.gpu

TOL = 1e-12

SPECS = {
    "quartic": lambda: specs.quartic(),
    "quartic27": lambda: specs.quartic(ndim=3, natom=9),
    "coupled2d": lambda: specs.coupled2d(),
    "evb": lambda: specs.evb(),
    "water2": lambda: specs.water_cluster(nmol=2),
    "water6": lambda: specs.water_cluster(nmol=6),
}


@pytest.mark.parametrize("name", list(SPECS))
def test_pes_potforce(name):
    sp = SPECS[name]()
    pes = Pes(sp)
    pes.set_V0(0.0123)
    op = oracle_pes(sp, 0.0123)
    x = rand_chain(sp, 37, seed=1, scale=0.05)
    g, e = pes.potforce_beads(x)
    for b in range(x.shape[1]):
        go, eo = op.potforce(x[:, b])
        assert abs(e[b] - eo) <= TOL * max(abs(eo), 1e-3)
        assert rel(g[:, b], go) <= TOL
    assert abs(pes.V(sp.well1) - op.V(sp.well1)) <= TOL * 1e-2


@pytest.mark.parametrize("name", list(SPECS))
def test_pes_hessian(name):
    sp = SPECS[name]()
    pes = Pes(sp)
    op = oracle_pes(sp)
    x = rand_chain(sp, 5, seed=2, scale=0.05)
    H = pes.Vdoubleprime(x)
    # the oracle is analytic except for WATER, where it is central FD with eps=1e-4
    tol = 1e-6 if sp.kind == specs.PES_WATER else 1e-11
    for b in range(x.shape[1]):
        Ho = op.Vdoubleprime(x[:, b])
        assert rel(H[b], Ho) <= tol
        assert rel(H[b], H[b].T) <= 1e-12


def test_water_param_grad():
    sp = specs.water_cluster(nmol=6)
    pes = Pes(sp)
    op = oracle_pes(sp)
    x = rand_chain(sp, 4, seed=3, scale=0.05)
    pg = pes.param_grad(x)
    for b in range(4):
        assert rel(pg[b], op.param_grad(x[:, b])) <= TOL


def _pair(spec, n, betan=30.0, tau=15.0, dt=1.0, gamma=1e-3, cayley=False):
    pes = Pes(spec)
    a, b = spec.well1, spec.well2
    ch = Chain(pes, n, spec.mass, betan, tau, dt, gamma, a, b, cayley=cayley)
    nm = O.NM(n, spec.ndim, spec.natom, spec.mass, betan, tau, dt, gamma, cayley, a, b)
    return pes, ch, nm, oracle_pes(spec)


@pytest.mark.parametrize("n", [16, 64, 100, 128])
def test_init_nm_and_transforms(n):
    sp = specs.evb()
    _, ch, nm, _ = _pair(sp, n)
    T, lam, bm, bv = ch.nm_arrays()
    assert np.array_equal(T, nm.T) and np.array_equal(lam, nm.lam)
    assert np.array_equal(bm, nm.beadmass) and rel(bv, nm.beadvec) <= 1e-15
    x = rand_chain(sp, n, seed=4)
    # nmtransform_forward with the beadvec shift (positions) and without (momenta)
    q = ch.nmtransform_forward(x, bead=np.arange(sp.ndof) + 1)
    qo = np.array([nm.forward(x[d], d) for d in range(sp.ndof)])
    assert rel(q, qo) <= TOL
    pt = ch.nmtransform_forward(x, bead=0)
    pto = np.array([nm.forward(x[d], -1) for d in range(sp.ndof)])
    assert rel(pt, pto) <= TOL
    xb = ch.nmtransform_backward(q, bead=np.arange(sp.ndof) + 1)
    xbo = np.array([nm.backward(qo[d], d) for d in range(sp.ndof)])
    assert rel(xb, xbo) <= TOL
    assert rel(xb, x) <= 1e-12  # T^2 = I round trip


def _state(sp, n, seed):
    rng = np.random.default_rng(seed)
    x = rand_chain(sp, n, seed)
    p = rng.standard_normal(x.shape) * np.sqrt(sp.mass.repeat(sp.ndim))[:, None] * 0.01
    return x, p


@pytest.mark.parametrize("name,n", [("quartic", 64), ("coupled2d", 128), ("evb", 32),
                                    ("evb", 96), ("water6", 32)])
def test_single_pieces(name, n):
    sp = SPECS[name]()
    pes, ch, nm, op = _pair(sp, n)
    x, p = _state(sp, n, 5)
    ch.set_state(x, p)
    ch.step_v(0.7)
    xo, po = nm.step_v(op, 0.7, x, p)
    xg, pg = ch.get_state()
    assert rel(pg, po) <= TOL and np.array_equal(xg, x)
    ch.set_state(x, p)
    ch.step_nm(0.5)
    xo, po = nm.step_nm(0.5, x, p)
    xg, pg = ch.get_state()
    assert rel(xg, xo) <= TOL and rel(pg, po) <= TOL
    noise = np.random.default_rng(6).standard_normal(n * sp.ndof)
    ch.set_state(x, p)
    ch.step_langevin(noise)
    po = nm.step_langevin(O.Rng(normals=noise), p)
    _, pg = ch.get_state()
    assert rel(pg, po) <= TOL


@pytest.mark.parametrize("cayley", [False, True])
def test_time_step_pile_injected(cayley):
    sp = specs.evb()
    n = 64
    pes, ch, nm, op = _pair(sp, n, cayley=cayley)
    x, p = _state(sp, n, 7)
    noise = np.random.default_rng(8).standard_normal(n * sp.ndof)
    ch.set_state(x, p)
    ch.time_step_pile(noise)
    xo, po = nm.time_step_pile(op, O.Rng(normals=noise), x, p)
    xg, pg = ch.get_state()
    assert rel(xg, xo) <= TOL and rel(pg, po) <= TOL


def test_time_step_nm():
    sp = specs.coupled2d()
    n = 128
    pes, ch, nm, op = _pair(sp, n)
    x, p = _state(sp, n, 9)
    ch.set_state(x, p)
    ch.time_step_nm()
    xo, po = nm.time_step_nm(op, x, p)
    xg, pg = ch.get_state()
    assert rel(xg, xo) <= TOL and rel(pg, po) <= TOL


@pytest.mark.parametrize("name,n,steps", [("quartic", 64, 200), ("evb", 32, 100),
                                          ("quartic", 300, 40)])
def test_propagate_pile_injected(name, n, steps):
    """The loop runs on the fused step kernel up to 256 beads and as per-step launch
    sequences beyond (n = 300, at betan = beta/(n+1) with beta = 2000: betan = 30 at that n
    diverges in the oracle too)."""
    sp = SPECS[name]()
    pes, ch, nm, op = _pair(sp, n, **({"betan": 2000.0 / (n + 1)} if n > 256 else {}))
    x, p = _state(sp, n, 10)
    dbdl = sp.well2 - sp.well1
    noise = np.random.default_rng(11).standard_normal(steps * n * sp.ndof)
    ch.set_state(x, p)
    dh, nre = ch.propagate_pimd_pile(steps, dbdl, imin=10, noise=noise)
    xo, po, dho, nreo = O.propagate_pile(nm, op, O.Rng(normals=noise), x, p, steps, dbdl, imin=10)
    xg, pg = ch.get_state()
    assert rel(xg, xo) <= 1e-10 and rel(pg, po) <= 1e-10
    assert abs(dh - dho) <= 1e-10 * abs(dho)


def test_propagate_pile_philox():
    sp = specs.evb()
    n = 32
    pes, ch, nm, op = _pair(sp, n)
    x, p = _state(sp, n, 12)
    dbdl = sp.well2 - sp.well1
    ch.set_state(x, p)
    ch.set_rng(1234, 77)
    dh, _ = ch.propagate_pimd_pile(50, dbdl)
    xo, po, dho, _ = O.propagate_pile(nm, op, O.Rng(seed=1234, stream=77), x, p, 50, dbdl)
    xg, pg = ch.get_state()
    assert ch.cursor == 50 * n * sp.ndof
    assert rel(xg, xo) <= 1e-10 and abs(dh - dho) <= 1e-10 * abs(dho)


@pytest.mark.parametrize("n", [64, 300])
def test_propagate_andersen_injected(n):
    sp = specs.quartic()
    pes, ch, nm, op = _pair(sp, n, betan=2000.0 / (n + 1), tau=15.0)
    x, p = _state(sp, n, 13)
    dbdl = sp.well2 - sp.well1
    kicks = np.array([3, 5, 2, 7, 4, 6, 3, 5, 2, 8, 4, 4, 9, 3, 5], dtype=np.int32)
    noise = np.random.default_rng(14).standard_normal(len(kicks) * n * sp.ndof)
    ch.set_state(x, p)
    dh = ch.propagate_pimd_nm(40, dbdl, noutput=5, imin=4, noise=noise, kicks=kicks)
    xo, po, dho = O.propagate_andersen(nm, op, O.Rng(normals=noise, kicks=kicks), x, p, 40,
                                       dbdl, 5, imin=4)
    xg, pg = ch.get_state()
    assert rel(xg, xo) <= 1e-10 and abs(dh - dho) <= 1e-10 * abs(dho)


def test_init_path_and_reinit():
    sp = specs.evb()
    n = 32
    pes, ch, nm, op = _pair(sp, n)
    lampath, path = specs.linear_path(sp.well1, sp.well2, npath=7)
    path = path + 0.02 * np.sin(np.arange(7))[None, :]  # curved: exercises the spline
    opath = O.Path(lampath, path)
    ch.set_path(lampath, path)
    noise = np.random.default_rng(15).standard_normal(200 * n * sp.ndof)
    ch.init_path(0.6, noise[: n * sp.ndof])
    xo, po = O.init_path(nm, opath, O.Rng(normals=noise[: n * sp.ndof]), 0.6)
    xg, pg = ch.get_state()
    assert rel(xg, xo) <= TOL and rel(pg, po) <= TOL
    # dHdrlimit small enough to force re-initialisations mid-trajectory
    dbdl = sp.well2 - sp.well1
    dh, nre = ch.propagate_pimd_pile(60, dbdl, dhdrlimit=300.0, xi=0.6, noise=noise[n * sp.ndof:])
    xo2, po2, dho, nreo = O.propagate_pile(nm, op, O.Rng(normals=noise[n * sp.ndof:]), xo, po, 60,
                                           dbdl, dhdrlimit=300.0, xi=0.6, path=opath)
    assert nre == nreo and nre > 0
    xg, pg = ch.get_state()
    assert rel(xg, xo2) <= 1e-10 and abs(dh - dho) <= 1e-10 * max(abs(dho), 1e-30)
