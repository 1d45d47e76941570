"""Pins of the oracle's PIMD half against what the mathematics fixes.

verletmodule.f90 cannot be compiled here (it includes Intel's mkl_vsl.fi, which the
image lacks), so its restatement in oracle/oracle.c is held to closed forms, library
routines and invariants instead, each chosen so that a plausible slip (a dropped term,
a wrong sign, index or transpose) fails:

* init_nm (:306-338): T against scipy's orthonormal DST-I, T symmetric with T^2 = I,
  (lambda_i betan)^2 against eigvalsh of tridiag(-1, 2, -1), beadvec = T x_eq with x_eq
  solved by numpy from L x_eq = a e_1 + b e_N.
* step_nm (:494-561): the equal-frequency Cartesian form of the exact rotation (no T
  in it) and its Cayley variant; time reversibility.
* step_langevin (:634-667): deterministic part c1^2 p~ with c1 = exp(-gamma dt / tau);
  noise amplitude sqrt(m'/betan (1 - c1^4)) from OU stationarity, per noise slot.
* step_v (:565-585), PES: Vprime and Vdoubleprime against central differences of V.
* time_step_pile / time_step_nm (:423-435, :291-302): composition order of the
  pinned pieces; bounded energy error of the symplectic splitting.
* propagate_pimd_pile / _nm with the TI estimator (:372-416, :190-250): the sampled
  mean of sum m (-x_N) db/dxi against the closed-form Gaussian-chain mean for a
  harmonic potential.
* gauleg (:124-160) against numpy's leggauss; spline/splint/splin_grad
  (instantonmod.f90:397-561) against scipy's natural CubicSpline.
* Philox4x32-10 against the Random123 known-answer vectors; the normal and Poisson
  draws against their distributions.
* UMhessian (instantonmod.f90:155-217): the free chain's spectrum in closed form with
  the missing bead 1-2 coupling.
* The RPI splitting chain (rpi_ser.f90:221-381) of a 1D double well against the
  splitting of an exact grid diagonalisation.
"""
import ctypes as C
import math

import numpy as np
import pytest
import scipy.fft
import scipy.interpolate
import scipy.linalg
import scipy.optimize as so
import scipy.stats

import oracle as O
from helpers import rel, spec_by_name

PI = O.OR_PI


def _lap(n):
    return 2.0 * np.eye(n) - np.eye(n, k=1) - np.eye(n, k=-1)


@pytest.mark.parametrize("n", [1, 2, 7, 32, 63])
def test_nm_matrix_is_orthonormal_dst1(n):
    nm = O.NM(n, 1, 1, [1836.0], 10.0, 5.0)
    T = nm.T
    assert np.array_equal(T, T.T)
    assert np.abs(T @ T - np.eye(n)).max() <= 1e-13
    dst = scipy.fft.dst(np.eye(n), type=1, norm="ortho", axis=0)
    assert np.abs(T - dst).max() <= 1e-13


@pytest.mark.parametrize("n,betan,tau", [(5, 30.0, 1.0), (40, 2000.0 / 41, 20.0)])
def test_nm_frequencies_masses_and_beadvec(n, betan, tau):
    mass = np.array([1836.0, 1000.0])
    a = np.array([0.3, -0.2, 0.1, 0.7, -1.0, 2.0])
    b = np.array([-0.4, 0.5, 0.2, 0.9, 1.0, -2.0])
    nm = O.NM(n, 3, 2, mass, betan, tau, a=a, b=b)
    ev = np.linalg.eigvalsh(_lap(n))
    assert np.abs((nm.lam * betan) ** 2 - ev).max() <= 1e-13 * ev.max()
    assert np.all(np.diff(nm.lam) > 0)
    assert rel(nm.beadmass, mass[None, :] * (nm.lam[:, None] * tau) ** 2) <= 1e-15
    # beadvec = T x_eq, x_eq solved independently from L x_eq = a e_1 + b e_N
    rhs = np.zeros((6, n))
    rhs[:, 0] += a
    rhs[:, -1] += b
    xeq = np.linalg.solve(_lap(n), rhs.T).T
    assert rel(nm.beadvec, xeq @ nm.T) <= 1e-12
    assert rel(xeq[:, -1], a + (b - a) * n / (n + 1)) <= 1e-13
    # forward / backward
    x = np.random.default_rng(1).standard_normal((6, n))
    for d in range(6):
        q = nm.forward(x[d], d)
        assert rel(q, np.linalg.solve(nm.T, x[d] - xeq[d])) <= 1e-12
        assert rel(nm.backward(q, d), x[d]) <= 1e-13
    # momenta: no shift
    assert rel(nm.forward(x[0], -1), nm.T @ x[0]) <= 1e-14


@pytest.mark.parametrize("cayley", [False, True])
def test_step_nm_equal_frequency_form(cayley):
    """A-step in Cartesian form without T: with w = h / tau,
    x' = x_eq + C (x - x_eq) + S betan^2/(m tau) L^-1 p,  p' = C p - S m tau/betan^2 L (x - x_eq),
    C, S = cos w, sin w (exact) or (4 - w^2)/(4 + w^2), 4w/(4 + w^2) (Cayley)."""
    n, betan, tau, h = 24, 30.0, 7.0, 2.5
    mass = np.array([1500.0, 2500.0])
    a, b = np.array([0.1, -0.3]), np.array([0.6, 0.4])
    nm = O.NM(n, 1, 2, mass, betan, tau, cayley=cayley, a=a, b=b)
    rng = np.random.default_rng(2)
    x = rng.standard_normal((2, n))
    p = 30.0 * rng.standard_normal((2, n))
    x1, p1 = nm.step_nm(h, x, p)
    w = h / tau
    Cc, S = (((4 - w * w) / (4 + w * w), 4 * w / (4 + w * w)) if cayley else (math.cos(w), math.sin(w)))
    L = _lap(n)
    for d in range(2):
        rhs = np.zeros(n)
        rhs[0], rhs[-1] = a[d], b[d]
        xeq = np.linalg.solve(L, rhs)
        m = mass[d]
        xe = xeq + Cc * (x[d] - xeq) + S * betan ** 2 / (m * tau) * np.linalg.solve(L, p[d])
        pe = Cc * p[d] - S * m * tau / betan ** 2 * (L @ (x[d] - xeq))
        assert rel(x1[d], xe) <= 1e-11
        assert rel(p1[d], pe) <= 1e-11
    if not cayley:
        x2, p2 = nm.step_nm(-h, x1, p1)
        assert rel(x2, x) <= 1e-12 and rel(p2, p) <= 1e-12


def test_step_langevin_decay_and_amplitude():
    n, betan, tau, dt, gamma = 6, 20.0, 5.0, 1.0, 0.3
    mass = np.array([1836.0, 12.0])
    nm = O.NM(n, 2, 2, mass, betan, tau, dt=dt, gamma=gamma)
    D = 4
    c1 = math.exp(-gamma * dt / tau)
    assert np.abs(nm.c1 - c1).max() <= 1e-15
    T = nm.T
    p = np.random.default_rng(3).standard_normal((D, n))
    # zero noise: p~ decays by c1^2
    out = nm.step_langevin(O.Rng(normals=np.zeros(n * D)), p)
    assert rel(out, c1 * c1 * p) <= 1e-13
    # one unit normal in slot (dofi, k): the mode k of dof dofi gets the OU amplitude
    for dofi in range(D):
        atom = dofi // 2
        for k in (0, n - 1):
            z = np.zeros(n * D)
            z[dofi * n + k] = 1.0
            out = nm.step_langevin(O.Rng(normals=z), np.zeros((D, n)))
            mprime = mass[atom] * (nm.lam[k] * tau) ** 2
            want = np.zeros((D, n))
            want[dofi] = T[:, k] * math.sqrt(mprime / betan * (1 - c1 ** 4))
            assert rel(out, want) <= 1e-13


@pytest.mark.parametrize("kind", ["quartic", "coupled2d", "evb", "water", "harmonic"])
def test_pes_gradients_and_hessians_against_differences(kind):
    sp = spec_by_name(kind)
    pes = O.Pes(sp.kind, sp.ndim, sp.natom, sp.dp, sp.ip, 0.0)
    x = (sp.well1 + sp.well2) / 2 + 0.05 * np.random.default_rng(4).standard_normal(sp.ndof)
    g = pes.Vprime(x)
    gf, e = pes.potforce(x)
    assert rel(gf, g) == 0.0 and e == pes.V(x)
    eps = 1e-5
    fd = np.array([(pes.V(x + eps * np.eye(sp.ndof)[d]) - pes.V(x - eps * np.eye(sp.ndof)[d])) / (2 * eps)
                   for d in range(sp.ndof)])
    assert rel(g, fd) <= 1e-6
    if kind != "water":  # the water Hessian is itself central differences of Vprime
        H = pes.Vdoubleprime(x)
        Hfd = np.array([(pes.Vprime(x + eps * np.eye(sp.ndof)[d]) - pes.Vprime(x - eps * np.eye(sp.ndof)[d])) / (2 * eps)
                        for d in range(sp.ndof)])
        assert rel(H, Hfd) <= 1e-6
        assert np.abs(H - H.T).max() <= 1e-12 * np.abs(H).max()


def test_step_v_is_a_gradient_kick():
    sp = spec_by_name("coupled2d")
    pes = O.Pes(sp.kind, sp.ndim, sp.natom, sp.dp, sp.ip, 0.0)
    n = 5
    nm = O.NM(n, 2, 1, sp.mass, 10.0, 3.0)
    x = np.random.default_rng(5).standard_normal((2, n)) * 0.3
    p = np.random.default_rng(6).standard_normal((2, n))
    x1, p1 = nm.step_v(pes, 0.7, x, p)
    want = p - 0.7 * np.stack([pes.Vprime(x[:, i]) for i in range(n)], axis=1)
    assert np.array_equal(x1, x) and rel(p1, want) <= 1e-15


def test_step_compositions():
    sp = spec_by_name("quartic")
    pes = O.Pes(sp.kind, sp.ndim, sp.natom, sp.dp, sp.ip, 0.0)
    n, dt = 9, 1.0
    nm = O.NM(n, 1, 1, sp.mass, 15.0, 10.0, dt=dt, gamma=0.1, a=sp.well1, b=sp.well2)
    x = np.linspace(-0.7, 0.7, n)[None, :] + 0.01
    p = np.random.default_rng(7).standard_normal((1, n))
    z = np.random.default_rng(8).standard_normal(n)
    x1, p1 = nm.time_step_pile(pes, O.Rng(normals=z), x, p)
    xa, pa = nm.step_v(pes, dt, x, p)
    xa, pa = nm.step_nm(dt / 2, xa, pa)
    pa = nm.step_langevin(O.Rng(normals=z), pa)
    xa, pa = nm.step_nm(dt / 2, xa, pa)
    assert rel(x1, xa) <= 1e-15 and rel(p1, pa) <= 1e-15
    x2, p2 = nm.time_step_nm(pes, x, p)
    xb, pb = nm.step_nm(dt / 2, x, p)
    xb, pb = nm.step_v(pes, dt, xb, pb)
    xb, pb = nm.step_nm(dt / 2, xb, pb)
    assert rel(x2, xb) <= 1e-15 and rel(p2, pb) <= 1e-15


def test_nve_energy_error_is_bounded():
    """gamma = 0: the A(dt/2) B(dt) A(dt/2) splitting of H = sum p~^2/2m' + ring springs +
    sum V (with m' the NM masses) keeps the energy error O(dt^2) and non-growing."""
    sp = spec_by_name("quartic")
    pes = O.Pes(sp.kind, sp.ndim, sp.natom, sp.dp, sp.ip, 0.0)
    n, betan, tau = 16, 2000.0 / 17, 20.0
    nm = O.NM(n, 1, 1, sp.mass, betan, tau, dt=1.0, a=sp.well1, b=sp.well2)
    m = sp.mass[0]
    T, mp = nm.T, nm.beadmass[:, 0]
    bv = nm.beadvec[0]

    def energy(x, p):
        q = T @ x[0] - bv
        pt = T @ p[0]
        # equal frequencies: H_free = sum p~^2/2m' + m' q^2 / (2 tau^2)
        return float(np.sum(pt ** 2 / (2 * mp) + mp * q ** 2 / (2 * tau ** 2)) +
                     sum(pes.V(x[:, i]) for i in range(n)))

    rng = np.random.default_rng(9)
    x = np.linspace(-0.7, 0.7, n)[None, :] + 0.02 * rng.standard_normal((1, n))
    p = (T @ (np.sqrt(mp / betan) * rng.standard_normal(n)))[None, :]
    e0 = energy(x, p)
    errs = []
    for _ in range(4000):
        x, p = nm.time_step_nm(pes, x, p)
        errs.append(energy(x, p) - e0)
    errs = np.abs(np.array(errs))
    assert errs.max() <= 2e-3 * abs(e0)
    assert errs[-1000:].max() <= 2.0 * errs[:1000].max() + 1e-12


def _harmonic_setup(n=8, beta=1000.0, kspring=0.0459, m=1836.0, a=-0.3, b=0.5):
    betan = beta / (n + 1)
    pes = O.Pes(O.PES_HARMONIC, 1, 1, [kspring, 0.0])
    A = (m / betan ** 2) * _lap(n) + kspring * np.eye(n)
    rhs = np.zeros(n)
    rhs[0], rhs[-1] = m * a / betan ** 2, m * b / betan ** 2
    xbar = np.linalg.solve(A, rhs)
    return pes, betan, xbar, m, np.array([a]), np.array([b])


@pytest.mark.parametrize("thermostat", ["pile", "andersen"])
def test_ti_estimator_harmonic_closed_form(thermostat):
    """<sum m (-x_N) db/dxi> in the Gaussian chain exp(-betan U) with U = sum k x^2/2 +
    springs m/(2 betan^2) incl. the fixed ends: the mean chain solves
    (m/betan^2 L + k) xbar = m/betan^2 (a e_1 + b e_N)."""
    n = 8
    pes, betan, xbar, m, a, b = _harmonic_setup(n)
    dbdl = np.array([1.0])
    want = -m * xbar[-1] * dbdl[0]
    means = []
    for traj in range(24):
        nm = O.NM(n, 1, 1, [m], betan, 20.0, dt=1.0, gamma=0.02, a=a, b=b)
        rng = O.Rng(seed=1234, stream=traj)
        x0 = xbar[None, :].copy()
        p0 = np.zeros((1, n))
        if thermostat == "pile":
            _, _, dh, _ = O.propagate_pile(nm, pes, rng, x0, p0, 60000, dbdl, imin=5000)
        else:
            _, _, dh = O.propagate_andersen(nm, pes, rng, x0, p0, 60000, dbdl, 40, imin=5000)
        means.append(dh)
    means = np.array(means)
    err = means.std(ddof=1) / math.sqrt(len(means))
    assert abs(means.mean() - want) <= 4.0 * err + 1e-3 * abs(want), (means.mean(), want, err)
    assert err <= 0.02 * abs(want)  # the test has power


def test_gauleg_against_numpy():
    for n in (1, 2, 5, 11, 16, 40):
        x, w = O.gauleg(0.0, 1.0, n)
        xr, wr = np.polynomial.legendre.leggauss(n)
        assert np.abs(x - (xr + 1) / 2).max() <= 1e-14
        assert np.abs(w - wr / 2).max() <= 1e-14


def test_spline_against_scipy_natural():
    rng = np.random.default_rng(10)
    lam = np.sort(np.concatenate([[0.0, 1.0], rng.uniform(0, 1, 7)]))
    path = rng.standard_normal((3, lam.size))
    pth = O.Path(lam, path)
    for xi in np.concatenate([lam, rng.uniform(0, 1, 20)]):
        cs = scipy.interpolate.CubicSpline(lam, path, axis=1, bc_type="natural")
        assert rel(pth.splint(xi), cs(xi)) <= 1e-12
        assert rel(pth.splin_grad(xi), cs(xi, 1)) <= 1e-11


def test_philox_known_answers():
    """Random123 kat_vectors for philox4x32-10."""
    kat = [((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
           ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2, (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
           ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
            (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1))]
    for ctr, key, want in kat:
        c = (C.c_uint32 * 4)(*ctr)
        k = (C.c_uint32 * 2)(*key)
        o = (C.c_uint32 * 4)()
        O.lib().or_philox4x32_10(c, k, o)
        assert tuple(o) == want


def test_philox_normals_and_poisson_distributions():
    z = O.philox_normals(77, 5, 0, 40000)
    assert scipy.stats.kstest(z, "norm").pvalue > 1e-3
    assert abs(z.mean()) < 4 / math.sqrt(z.size) and abs(z.var() - 1) < 0.03
    for mean in (3.0, 40.0):
        k = np.array([O.philox_poisson(9, 2, i, mean) for i in range(20000)])
        assert abs(k.mean() - mean) < 4 * math.sqrt(mean / k.size)
        assert abs(k.var() / mean - 1) < 0.05
        if mean < 16:
            vals, cnt = np.unique(k, return_counts=True)
            exp = scipy.stats.poisson.pmf(vals, mean) * k.size
            sel = exp > 20
            chi2 = np.sum((cnt[sel] - exp[sel]) ** 2 / exp[sel])
            assert scipy.stats.chi2.sf(chi2, sel.sum() - 1) > 1e-4


def test_free_chain_band_spectrum_with_the_quirk():
    """V = 0: bead 1 is decoupled (no 1-2 block, instantonmod.f90:203), so the spectrum is
    2/betan^2 once plus (2/betan^2)(1 - cos(k pi / N)), k = 1..N-1, per dof."""
    n, betan = 12, 7.0
    pes = O.Pes(O.PES_HARMONIC, 2, 1, [0.0, 0.0, 0.0, 0.0])
    ch = O.Chain(n, 2, 1, [1836.0], betan, True)
    ev = O.detj_eigs(ch.UMhessian(pes, np.zeros((2, n)), False))
    k = np.arange(1, n)
    want = np.sort(np.concatenate([[2 / betan ** 2] * 2, np.repeat(2 / betan ** 2 * (1 + np.cos(k * np.pi / n)), 2)]))
    assert np.abs(ev - want).max() <= 1e-13


def test_rpi_splitting_of_a_double_well_against_grid_diagonalisation():
    """The whole RPI chain of the oracle (UM + gradient, minimisation, band, lndetJ0/lndetJ,
    theta) for V0 ((x/a)^2 - 1)^2 against E1 - E0 of a fine finite-difference grid."""
    V0, a, m = 0.02, 1.0, 1836.0
    pes = O.Pes(O.PES_QUARTIC, 1, 1, [V0, a])
    # exact splitting: 3-point Laplacian on a fine grid
    x = np.linspace(-2.6, 2.6, 6001)
    hgrid = x[1] - x[0]
    d = 1.0 / (m * hgrid * hgrid) + V0 * ((x / a) ** 2 - 1) ** 2
    e = np.full(x.size - 1, -0.5 / (m * hgrid * hgrid))
    E = scipy.linalg.eigh_tridiagonal(d, e, select="i", select_range=(0, 1), eigvals_only=True)
    exact = E[1] - E[0]
    beta, n = 12000.0, 256
    betan = beta / n
    ch = O.Chain(n, 1, 1, [m], betan, True)
    wa, wb = np.array([-a]), np.array([a])
    x0 = np.linspace(-a, a, n)[None, :] + 0.0

    def fg(v):
        g, e_ = ch.UMforceenergy(pes, v.reshape(1, n), wa, wb)
        return e_, g.reshape(-1)

    res = so.minimize(fg, x0.reshape(-1), jac=True, method="L-BFGS-B",
                      options={"maxcor": 8, "ftol": 1e-15, "gtol": 1e-11, "maxiter": 50000})
    xi = res.x.reshape(1, n)
    ev0 = O.detj_eigs(ch.UMhessian(pes, np.full((1, n), -a), True))
    ev = O.detj_eigs(ch.UMhessian(pes, xi, False))
    assert abs(ev[0]) < 1e-3 * ev[1]  # the zero mode of the kink
    _, _, delta = O.rpi_splitting(betan, betan * res.fun, O.lndetj(ev), O.lndetj0(ev0))
    assert abs(delta / exact - 1) < 0.1, (delta, exact)
