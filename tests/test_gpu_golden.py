"""The device path against the compiled reference's own outputs (tests/golden/ref_*.npz,
from instantonmod.f90's UM, UMforceenergy, UMprime, UMhessian and instanton).

Bars as for the oracle: 1e-12 for U_M and gradients, 1e-12 / 1e-9 (analytic / the
reference plugin's finite-difference EVB Hessian) for the band; the instanton action
at or below the reference's L-BFGS-B result and within its stopping slack."""
import os

import numpy as np
import pytest

from helpers import rel
from rptunnel import Pes, Ring
from rptunnel.specs import PesSpec

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(name):
    f = dict(np.load(os.path.join(GOLD, name)))
    ip = f["ip"].astype(np.int32) if f["ip"].size else None
    sp = PesSpec("golden", int(f["kind"]), int(f["ndim"]), int(f["natom"]), f["dp"], ip,
                 f["mass"], f["well1"], f["well2"])
    return f, sp


@pytest.mark.parametrize("name", ["quartic", "coupled2d", "harmonic", "evb"])
def test_device_um_and_band_against_reference(name):
    f, sp = _load(f"ref_um_{name}.npz")
    pes = Pes(sp)
    pes.set_V0(float(f["v0"]))
    n = int(f["n"])
    for tag, fixed in (("fixed", True), ("free", False)):
        ring = Ring(pes, sp.mass, [n], [float(f["betan"])], sp.well1, sp.well2, fixedends=fixed)
        g, e = ring.UMforceenergy([f["x"]])
        assert abs(e[0] - f[f"ufe_{tag}"]) <= 1e-12 * abs(f[f"ufe_{tag}"])
        assert rel(g[0], f[f"grad_{tag}"]) <= 1e-12
        assert rel(g[0], f[f"prime_{tag}"]) <= 1e-12
    ring = Ring(pes, sp.mass, [n], [float(f["betan"])], sp.well1, sp.well2)
    ring.set_x([f["x"]])
    tol = 1e-9 if name == "evb" else 1e-12
    for sw in (0, 1):
        assert rel(ring.UMhessian(0, bool(sw)), f[f"band_sw{sw}"]) <= tol


@pytest.mark.parametrize("name", ["quartic", "coupled2d", "evb"])
def test_device_instanton_against_reference(name):
    f, sp = _load(f"ref_instanton_{name}.npz")
    pes = Pes(sp)
    n = int(f["n"])
    ring = Ring(pes, sp.mass, [n], [float(f["betan"])], sp.well1, sp.well2)
    act_ref = float(f["action"])
    # the reference's settings (instantonmod.f90:719-724): m = 8, factr = 1e6, pgtol = eps2
    _, act, _, st = ring.instanton([f["x0"]], pgtol=float(f["eps2"]), factr=1e6, m=8)
    assert st[0] == 0
    assert abs(act[0] - act_ref) <= 1e-5 * abs(act_ref)
    # tightly converged: at or below the reference's stopping point
    xo, act, _, st = ring.instanton([f["x0"]], pgtol=1e-10, factr=10.0, m=8)
    assert st[0] == 0
    assert act[0] <= act_ref + 1e-15 and act_ref - act[0] <= 1e-5 * abs(act_ref)
    assert rel(xo[0], f["xopt"]) <= 1e-2
