"""The batched TI workloads at their full sizes, on the launch shape they are timed with: the
27-dof EVB double well, N in {32, 64, 96, 128, 192, 256}, 512 trajectories x 11 lambda nodes
at every N (33,792 polymers); the 2D coupled double well, N = 128, 4096 trajectories x 16
lambda nodes (65,536 polymers); the 6-water cluster (54 dof), N = 128, 1024 trajectories x
11 lambda nodes (11,264 polymers).  Langevin, the engine's own Philox streams after
init_path, 8 steps per launch from captured graphs (propagate_pimd_pile,
verletmodule.f90:372-416).

Checked without shrinking it:
* the oracle replays the first and the last polymer of every bead-count family from the
  state init_path gave them and the stream position it left: states and TI averages 1e-10
  normwise relative, stream cursors exact;
* a second batch built the same way gives bit-identical TI sums, counts and cursors for every
  polymer, and bit-identical states for the sampled ones (one workgroup per polymer, no
  atomics: the run is deterministic);
* every TI sum is finite and every polymer counted the same number of steps.
"""
import numpy as np
import pytest

import oracle as O
from helpers import oracle_pes, rel
from rptunnel import Batch, Pes, specs, ti

pytestmark = pytest.mark.gpu

BETA, TAU, DT, GAMMA, SEED = 2000.0, 20.0, 1.0, 1e-3, 2024
STEPS, IMIN = 16, 4
# (potential, bead counts, trajectories per lambda node and bead count, lambda nodes,
# polymers)
WORKLOADS = [("evb27", [32, 64, 96, 128, 192, 256], 512, 11, 33792),
             ("coupled2d", [128], 4096, 16, 65536),
             ("water6", [128], 1024, 11, 11264)]


def _spec(name):
    return {"evb27": specs.evb, "coupled2d": specs.coupled2d,
            "water6": lambda: specs.water_cluster(nmol=6)}[name]()


def _path(name, sp):
    """(lampath, path (ndof, npath)): the straight line between the wells for the EVB; for
    the water cluster the path that turns molecule 0's H pair about its bisector (a straight
    line to the exchanged well would run the two H atoms through each other)."""
    if name != "water6":
        return specs.linear_path(sp.well1, sp.well2, npath=2)
    path = np.ascontiguousarray(sp.extra["path"].T)
    return specs.arc_lampath(path), path


def _polymers(sp, ns, ntraj, nlam, lampath, path):
    t, _ = np.polynomial.legendre.leggauss(nlam)
    xi = 0.5 * (t + 1.0)  # Gauss-Legendre nodes on [0, 1]
    ends, grads = ti._spline_nodes(lampath, path, xi)  # the path's point and slope at xi
    pols, stream = [], 0
    for n in ns:
        for il, x in enumerate(xi):
            for _ in range(ntraj):
                pols.append(dict(n=n, betan=BETA / (n + 1), xi=float(x), a=path[:, 0],
                                 b=ends[il], dbdl=grads[il], stream=stream))
                stream += 1
    return pols


def _batch(pes, sp, pols, lampath, path):
    b = Batch(pes, sp.mass, pols, TAU, DT, GAMMA, seed=SEED)
    b.set_path(lampath, path)
    b.init_path()
    b.set_launch(steps_per_launch=8)
    return b


@pytest.mark.parametrize("name,ns,ntraj,nlam,npoly", WORKLOADS)
def test_full_workload_against_oracle_and_deterministic(name, ns, ntraj, nlam, npoly):
    sp = _spec(name)
    pes = Pes(sp)
    op = oracle_pes(sp)
    lampath, path = _path(name, sp)
    pols = _polymers(sp, ns, ntraj, nlam, lampath, path)
    per = ntraj * nlam
    assert len(pols) == len(ns) * per == npoly
    sample = [k for f in range(len(ns)) for k in (f * per, (f + 1) * per - 1)]
    b = _batch(pes, sp, pols, lampath, path)
    start = {k: b.get_state(k) for k in sample}
    cur0 = b.cursors()
    b.run(STEPS, imin=IMIN)
    sums, cnt, _ = b.ti_readout()
    cur = b.cursors()
    assert np.all(cnt == STEPS - IMIN) and np.all(np.isfinite(sums))
    for k in sample:
        pol = pols[k]
        nm = O.NM(pol["n"], sp.ndim, sp.natom, sp.mass, pol["betan"], TAU, DT, GAMMA, False,
                  pol["a"], pol["b"])
        r = O.Rng(seed=SEED, stream=pol["stream"])
        r.s.cursor = int(cur0[k])
        x0, p0 = start[k]
        xo, po, dho, _ = O.propagate_pile(nm, op, r, x0, p0, STEPS, pol["dbdl"], imin=IMIN)
        xg, pg = b.get_state(k)
        assert cur[k] == r.cursor, (k, cur[k], r.cursor)
        assert rel(xg, xo) <= 1e-10, (k, pol["n"], rel(xg, xo))
        assert rel(pg, po) <= 1e-10, (k, pol["n"], rel(pg, po))
        assert abs(sums[k] / (STEPS - IMIN) - dho) <= 1e-10 * abs(dho), (k, sums[k], dho)
    states = {k: b.get_state(k) for k in sample}
    b.close()
    b2 = _batch(pes, sp, pols, lampath, path)
    b2.run(STEPS, imin=IMIN)
    sums2, cnt2, _ = b2.ti_readout()
    assert np.array_equal(sums2, sums) and np.array_equal(cnt2, cnt)
    assert np.array_equal(b2.cursors(), cur)
    for k in sample:
        x2, p2 = b2.get_state(k)
        assert np.array_equal(x2, states[k][0]) and np.array_equal(p2, states[k][1]), k
    b2.close()


def test_full_instanton_workload():
    """The batched-instanton workload at its full size: 256 EVB chains, N uniform in
    [64, 1024], beta in [1000, 8000] (betan = beta / N, rpi_ser.f90:45).  The spring-
    preconditioned search from straight lines converges every chain with the reference's
    stopping rule (m = 8, factr = 1e6, pgtol = 1e-5); at the instantons every Hessian is
    positive definite, lndetJ is finite, and for the longest chain and one of median length
    eta_1 and lndetJ match LAPACK's dsbevx (lowest eigenvalue of the band) and the banded
    Cholesky.  Bars: lndetJ 1e-9 relative; eta_1 1e-7 relative.  eta_1 is some 3e-7
    against a band norm near 0.1, and the device's harmonic Rayleigh quotient from inverse
    iteration shifted close below it agrees with dsbevx to about 2e-8 relative (6e-15
    absolute) at N = 1021; the bar still tells eta_1 from eta_2, which the test checks."""
    import scipy.linalg as sl
    from rptunnel import Ring

    sp = specs.evb()
    pes = Pes(sp)
    rng = np.random.default_rng(4)
    ns = [int(v) for v in rng.integers(64, 1025, 256)]
    betas = rng.uniform(1000.0, 8000.0, 256)
    betan = [bt / n for bt, n in zip(betas, ns)]
    ring = Ring(pes, sp.mass, ns, betan, sp.well1, sp.well2)
    x0 = []
    for n in ns:
        k = np.arange(1, n + 1)
        x0.append((np.outer(sp.well1, n - k) + np.outer(sp.well2, k - 1)) / (n - 1))
    pes.set_V0(pes.V(sp.well1))
    ring.set_preconditioner(0.05)
    xo, act, it, st = ring.instanton(x0, pgtol=1e-5, factr=1e6, m=8, maxiter=0)
    assert np.all(st == 0) and np.all(it <= 60), (np.unique(st), it.max())
    g, e = ring.UMforceenergy(xo)
    assert np.allclose(e, act, rtol=1e-12, atol=0.0)
    ring.set_x(xo)
    ld, nn, eta = ring.logdet(singlewell=False, skip_lowest=True)
    assert np.all(nn == 0) and np.all(eta > 0.0) and np.all(np.isfinite(ld))
    order = np.argsort(ns)
    for c in (int(order[-1]), int(order[len(ns) // 2])):
        band = ring.UMhessian(c, singlewell=False)
        ev = sl.eig_banded(band, lower=True, eigvals_only=True, select="i", select_range=(0, 1))
        L = sl.cholesky_banded(band, lower=True)
        want = 2.0 * float(np.sum(np.log(L[0]))) - float(np.log(ev[0]))
        assert ev[1] - ev[0] > 100 * 1e-7 * ev[0], (ns[c], ev)
        assert abs(eta[c] - ev[0]) <= 1e-7 * ev[0], (ns[c], eta[c], ev[0])
        assert abs(ld[c] - want) <= 1e-9 * abs(want), (ns[c], ld[c], want)
    pes.set_V0(0.0)
    ring.close()
