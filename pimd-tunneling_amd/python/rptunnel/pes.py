"""The ``mcmod_mass`` plugin API over the device potentials.

Mirrors mcmod_waterdimer.f90:12-103: ``V_init``, ``V``, ``Vprime``, ``potforce``,
``Vdoubleprime`` and the module variable ``V0``.  Every evaluation runs on the GPU
(there is no CPU path); one-bead calls are simply batches of one.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from .specs import PesSpec


class Pes:
    def __init__(self, spec: PesSpec):
        L.init()
        self.spec = spec
        self.ndim, self.natom, self.ndof = spec.ndim, spec.natom, spec.ndof
        dp = L.f64(spec.dp)
        ipa = None if spec.ip is None else np.ascontiguousarray(spec.ip, dtype=np.int32)
        h = C.c_void_p()
        L.check(L.lib().rp_pes_create(spec.kind, spec.ndim, spec.natom, L.ptr(dp), dp.size,
                                      L.iptr(ipa), 0 if ipa is None else ipa.size, C.byref(h)))
        self.h = h
        self.V0 = 0.0

    def close(self):
        if getattr(self, "h", None):
            L.lib().rp_pes_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- module variable V0 (pimd_par.f90:166 sets V0 = V(well1))
    def set_V0(self, v0):
        self.V0 = float(v0)
        L.check(L.lib().rp_pes_set_v0(self.h, self.V0))

    def V_init(self, iproc=0):
        """mcmod_waterdimer.f90:12-17: V0 = 0."""
        self.set_V0(0.0)

    def _beads(self, x):
        x = np.asarray(x, dtype=np.float64)
        if x.ndim == 1:
            return x.reshape(self.ndof, 1), 1
        return np.ascontiguousarray(x), x.shape[1]

    def potforce_beads(self, x):
        """x: (ndof, nbead) -> (grad (ndof, nbead), energy (nbead,))."""
        xb, nb = self._beads(x)
        g = np.zeros((self.ndof, nb))
        e = np.zeros(nb)
        L.check(L.lib().rp_pes_potforce(self.h, nb, L.ptr(xb), L.ptr(g), L.ptr(e)))
        return g, e

    def V(self, x):
        g, e = self.potforce_beads(x)
        return float(e[0]) if np.asarray(x).ndim == 1 else e

    def Vprime(self, x):
        g, e = self.potforce_beads(x)
        return g[:, 0] if np.asarray(x).ndim == 1 else g

    def potforce(self, x):
        g, e = self.potforce_beads(x)
        return (g[:, 0], float(e[0])) if np.asarray(x).ndim == 1 else (g, e)

    def Vdoubleprime(self, x):
        """hess[i, k] = d grad_k / d x_i (per bead: shape (nbead, ndof, ndof))."""
        xb, nb = self._beads(x)
        h = np.zeros((nb, self.ndof, self.ndof))
        L.check(L.lib().rp_pes_hessian(self.h, nb, L.ptr(xb), L.ptr(h)))
        return h[0] if np.asarray(x).ndim == 1 else h

    def param_grad(self, x):
        xb, nb = self._beads(x)
        out = np.zeros((nb, 4))
        npar = C.c_int(0)
        L.check(L.lib().rp_pes_param_grad(self.h, nb, L.ptr(xb), L.ptr(out), C.byref(npar)))
        out = out.reshape(-1)[: nb * npar.value].reshape(nb, npar.value)
        return out[0] if np.asarray(x).ndim == 1 else out
