"""The oracle against the compiled reference (instantonmod.f90 built from the reference's
own sources, see oracle/ref/build_ref.sh).

Committed fixtures (tests/golden/ref_*.npz, written by tests/golden/make_golden.py)
hold the reference's UM, UMforceenergy, UMprime and UMhessian outputs and its
L-BFGS-B instanton; when ``oracle/_ref/ref_instanton`` is built in this container the
same comparisons also run live on fresh random chains.

Bars: action and gradients 1e-13 relative (same formula, same FP64 order up to
compiler contraction); bands 1e-12 where both sides use analytic Hessians and 1e-9 for
the EVB (the reference plugin differentiates Vprime by central differences, eps = 1e-4,
mcmod_waterdimer.f90:83-103; the oracle's EVB Hessian is analytic).
"""
import os

import numpy as np
import pytest
import scipy.optimize as so

import oracle as O
import refrun as R
from helpers import rel

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
UM_CASES = ["quartic", "coupled2d", "harmonic", "evb"]
INST_CASES = ["quartic", "coupled2d", "evb"]


def _load(name):
    return dict(np.load(os.path.join(GOLD, name)))


def _pes(f, v0=0.0):
    ip = f["ip"] if f["ip"].size else None
    return O.Pes(int(f["kind"]), int(f["ndim"]), int(f["natom"]), f["dp"], ip, v0)


def _band_tol(kind):
    return 1e-9 if int(kind) == O.PES_EVB else 1e-12


@pytest.mark.parametrize("name", UM_CASES)
def test_um_against_reference_fixture(name):
    f = _load(f"ref_um_{name}.npz")
    op = _pes(f, float(f["v0"]))
    x, n = f["x"], int(f["n"])
    for tag, fixed in (("fixed", True), ("free", False)):
        oc = O.Chain(n, int(f["ndim"]), int(f["natom"]), f["mass"], float(f["betan"]), fixed)
        g, e = oc.UMforceenergy(op, x, f["well1"], f["well2"])
        assert abs(e - f[f"ufe_{tag}"]) <= 1e-13 * abs(f[f"ufe_{tag}"])
        assert abs(oc.UM(op, x, f["well1"], f["well2"]) - f[f"um_{tag}"]) <= 1e-13 * abs(f[f"um_{tag}"])
        assert rel(g, f[f"grad_{tag}"]) <= 1e-13
        assert rel(oc.UMprime(op, x, f["well1"], f["well2"]), f[f"prime_{tag}"]) <= 1e-13


@pytest.mark.parametrize("name", UM_CASES)
def test_umhessian_against_reference_fixture(name):
    f = _load(f"ref_um_{name}.npz")
    op = _pes(f, float(f["v0"]))
    n = int(f["n"])
    oc = O.Chain(n, int(f["ndim"]), int(f["natom"]), f["mass"], float(f["betan"]), True)
    for sw in (0, 1):
        band = oc.UMhessian(op, f["x"], bool(sw))
        assert rel(band, f[f"band_sw{sw}"]) <= _band_tol(f["kind"])
        # the reference's quirk survives in the fixture itself: no bead 1-2 coupling
        D = int(f["ndim"]) * int(f["natom"])
        assert np.all(f[f"band_sw{sw}"][D, :D] == 0.0)


@pytest.mark.parametrize("name", INST_CASES)
def test_instanton_against_reference_fixture(name):
    """The reference stops L-BFGS-B at factr = 1e6 / pgtol = 1e-5 (instantonmod.f90:723-724);
    the converged minimum is unique, so the oracle's U_M minimised tightly (scipy L-BFGS-B)
    sits at or below the reference's action, within the reference's stopping slack."""
    f = _load(f"ref_instanton_{name}.npz")
    op = _pes(f)
    n, D = int(f["n"]), int(f["ndim"]) * int(f["natom"])
    oc = O.Chain(n, int(f["ndim"]), int(f["natom"]), f["mass"], float(f["betan"]), True)
    # the oracle's action at the reference's optimum is the reference's action
    assert abs(oc.UM(op, f["xopt"], f["well1"], f["well2"]) - f["action"]) <= 1e-13 * abs(f["action"])
    g, _ = oc.UMforceenergy(op, f["xopt"], f["well1"], f["well2"])
    assert np.abs(g).max() <= 2e-5  # near the reference's pgtol

    def fg(v):
        gg, e = oc.UMforceenergy(op, v.reshape(D, n), f["well1"], f["well2"])
        return e, gg.reshape(-1)

    res = so.minimize(fg, f["x0"].reshape(-1), jac=True, method="L-BFGS-B",
                      options={"maxcor": 8, "ftol": 1e-15, "gtol": 1e-10, "maxiter": 20000})
    act = float(f["action"])
    assert res.fun <= act + 1e-15
    assert act - res.fun <= 1e-5 * abs(act)
    assert rel(res.x.reshape(D, n), f["xopt"]) <= 1e-2


needs_ref = pytest.mark.skipif(not R.available(), reason="oracle/_ref not built in this container")


@needs_ref
@pytest.mark.parametrize("name", UM_CASES)
@pytest.mark.parametrize("fixed", [True, False])
def test_um_live_reference(name, fixed):
    f = _load(f"ref_um_{name}.npz")
    rng = np.random.default_rng(41)
    D = int(f["ndim"]) * int(f["natom"])
    for n in (2, 5, 13):
        x = np.linspace(0, 1, n)[None, :] * (f["well2"] - f["well1"])[:, None] + f["well1"][:, None]
        x = x + 0.07 * rng.standard_normal((D, n))
        betan = float(rng.uniform(3.0, 40.0))
        ip = f["ip"] if f["ip"].size else None
        e, e2, g, g2 = R.um(int(f["kind"]), int(f["ndim"]), int(f["natom"]), f["dp"], ip,
                            f["mass"], betan, x, f["well1"], f["well2"], fixed, 0.002)
        oc = O.Chain(n, int(f["ndim"]), int(f["natom"]), f["mass"], betan, fixed)
        op = _pes(f, 0.002)
        go, eo = oc.UMforceenergy(op, x, f["well1"], f["well2"])
        assert abs(eo - e) <= 1e-13 * abs(e) and abs(eo - e2) <= 1e-13 * abs(e)
        assert rel(go, g) <= 1e-13 and rel(go, g2) <= 1e-13
        for sw in (False, True):
            bo = oc.UMhessian(op, x, sw)
            br = R.hessian(int(f["kind"]), int(f["ndim"]), int(f["natom"]), f["dp"], ip,
                           f["mass"], betan, x, sw, 0.002)
            assert rel(bo, br) <= _band_tol(f["kind"])
