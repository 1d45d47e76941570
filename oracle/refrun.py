"""Driver of the compiled reference -- TEST INFRASTRUCTURE ONLY.

``oracle/ref/build_ref.sh`` compiles the reference's own instantonmod.f90 (with its
vendored L-BFGS-B) against the ``mcmod_mass`` plugin of ``oracle/ref/ref_plugin.f90``
into ``oracle/_ref/ref_instanton``.  This module feeds that program a potential, a
chain and its end points over stdin and parses what it prints, so the tests can hold
the C oracle (and, through the committed fixtures, the device path) against the
reference's own UM, UMforceenergy, UMprime, UMhessian and instanton routines.
Only tests/ and tests/golden/make_golden.py use it.
"""
from __future__ import annotations

import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
EXE = os.path.join(_HERE, "_ref", "ref_instanton")


def available():
    return os.path.exists(EXE)


def _fmt(a):
    return " ".join(repr(float(v)) for v in np.asarray(a, dtype=np.float64).reshape(-1))


def _run(task, kind, ndim, natom, dp, ip, mass, betan, x, a, b, fixedends=True,
         singlewell=False, v0=0.0, eps2=1e-5):
    x = np.asarray(x, dtype=np.float64)
    D, n = x.shape
    ip = [] if ip is None else list(np.asarray(ip, dtype=np.int64))
    lines = [task,
             f"{kind} {ndim} {natom} {n} {len(dp)} {len(ip)}",
             _fmt(dp)]
    if ip:
        lines.append(" ".join(str(int(v)) for v in ip))
    lines += [_fmt(mass),
              f"{float(betan)!r} {int(bool(fixedends))} {int(bool(singlewell))} {float(v0)!r} {float(eps2)!r}",
              _fmt(a), _fmt(b),
              # x(n, ndim, natom) column-major has the memory order of numpy (ndof, n)
              _fmt(x.reshape(-1))]
    out = subprocess.run([EXE], input="\n".join(lines) + "\n", capture_output=True, text=True,
                         check=True).stdout
    return out.split("RESULT", 1)[1]


def _floats(text):
    return np.array([float(t.replace("D", "E")) for t in text.split()], dtype=np.float64)


def um(kind, ndim, natom, dp, ip, mass, betan, x, a, b, fixedends=True, v0=0.0):
    """-> (UM, UMforceenergy energy, UMforceenergy gradient, UMprime) of the reference."""
    D, n = np.shape(x)
    v = _floats(_run("um", kind, ndim, natom, dp, ip, mass, betan, x, a, b, fixedends, v0=v0))
    return v[0], v[1], v[2:2 + D * n].reshape(D, n), v[2 + D * n:2 + 2 * D * n].reshape(D, n)


def hessian(kind, ndim, natom, dp, ip, mass, betan, x, singlewell, v0=0.0):
    """-> the reference's UMhessian band, shape (ndof+1, n*ndof)."""
    D, n = np.shape(x)
    z = np.zeros(D)
    v = _floats(_run("hessian", kind, ndim, natom, dp, ip, mass, betan, x, z, z, True,
                     singlewell=singlewell, v0=v0))
    return v.reshape(n * D, D + 1).T


def instanton(kind, ndim, natom, dp, ip, mass, betan, x, a, b, eps2=1e-5, v0=0.0):
    """-> (U_M at the reference's L-BFGS-B optimum, optimised chain (ndof, n))."""
    D, n = np.shape(x)
    text = _run("instanton", kind, ndim, natom, dp, ip, mass, betan, x, a, b, True,
                v0=v0, eps2=eps2)
    v = _floats(text.split("OPTIMUM", 1)[1])
    return v[0], v[1:].reshape(D, n)
