"""Writes the reference fixtures under tests/golden/ from the compiled reference.

Run in a container that holds the reference sources:

    bash oracle/ref/build_ref.sh && python tests/golden/make_golden.py

``oracle/_ref/ref_instanton`` is the reference's own instantonmod.f90 (UM,
UMforceenergy, UMprime, UMhessian, instanton with its vendored L-BFGS-B v3.0) built
with ``-fdefault-real-8 -fdefault-double-8`` against the ``mcmod_mass`` plugin in
oracle/ref/ref_plugin.f90.  Each fixture holds its inputs (potential parameters,
masses, betan, end points, chains) and the reference's outputs, so the tests can
check the oracle and the device path without the reference present.
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "oracle"), os.path.join(ROOT, "pimd-tunneling_amd", "python")]

import refrun as R  # noqa: E402
from rptunnel import specs  # noqa: E402


def _spec_fields(sp):
    ip = np.zeros(0, np.int64) if sp.ip is None else np.asarray(sp.ip, np.int64)
    return dict(kind=sp.kind, ndim=sp.ndim, natom=sp.natom, dp=np.asarray(sp.dp, float), ip=ip,
                mass=np.asarray(sp.mass, float), well1=sp.well1, well2=sp.well2)


def _chain(sp, n, seed, scale):
    rng = np.random.default_rng(seed)
    t = (np.arange(n) + 1.0) / (n + 1.0)
    x = np.outer(sp.well1, 1 - t) + np.outer(sp.well2, t)
    return x + scale * rng.standard_normal(x.shape)


def _linear(sp, n):
    k = np.arange(1, n + 1)
    return (np.outer(sp.well1, n - k) + np.outer(sp.well2, k - 1)) / (n - 1)


CASES = {
    "quartic": (specs.quartic, 9),
    "coupled2d": (specs.coupled2d, 9),
    "harmonic": (lambda: specs.harmonic([1e-3, 2e-3], [0.1, -0.2], [1000.0, 2000.0]), 9),
    "evb": (specs.evb, 6),
}


def make_um(name):
    """UM / UMforceenergy / UMprime at fixed and free ends, UMhessian both ways."""
    sp, n = CASES[name][0](), CASES[name][1]
    betan, v0 = 9.0, 0.001
    x = _chain(sp, n, 3, 0.05)
    f = _spec_fields(sp)
    ip = None if sp.ip is None else sp.ip
    out = dict(f, n=n, betan=betan, v0=v0, x=x)
    for tag, fixed in (("fixed", True), ("free", False)):
        e, e2, g, g2 = R.um(sp.kind, sp.ndim, sp.natom, sp.dp, ip, sp.mass, betan, x,
                            sp.well1, sp.well2, fixed, v0)
        out[f"um_{tag}"], out[f"ufe_{tag}"], out[f"grad_{tag}"], out[f"prime_{tag}"] = e, e2, g, g2
    for sw in (False, True):
        out[f"band_sw{int(sw)}"] = R.hessian(sp.kind, sp.ndim, sp.natom, sp.dp, ip, sp.mass,
                                             betan, x, sw, v0)
    np.savez_compressed(os.path.join(HERE, f"ref_um_{name}.npz"), **out)


INST = {
    "quartic": (specs.quartic, 16, 3000.0),
    "coupled2d": (specs.coupled2d, 16, 3000.0),
    "evb": (specs.evb, 16, 3000.0),
}


def make_instanton(name):
    """The reference's instanton() from a straight line, RPI betan = beta/N (rpi_ser.f90:45)."""
    mk, n, beta = INST[name]
    sp = mk()
    betan = beta / n
    x0 = _linear(sp, n)
    act, xo = R.instanton(sp.kind, sp.ndim, sp.natom, sp.dp, sp.ip, sp.mass, betan, x0,
                          sp.well1, sp.well2, eps2=1e-5)
    np.savez_compressed(os.path.join(HERE, f"ref_instanton_{name}.npz"), **_spec_fields(sp),
                        n=n, beta=beta, betan=betan, eps2=1e-5, x0=x0, action=act, xopt=xo)


if __name__ == "__main__":
    if not R.available():
        sys.exit("oracle/_ref/ref_instanton is not built (run oracle/ref/build_ref.sh)")
    for name in CASES:
        make_um(name)
    for name in INST:
        make_instanton(name)
    print("wrote", sorted(f for f in os.listdir(HERE) if f.endswith(".npz")))
