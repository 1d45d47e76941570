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
    """A device potential (``spec``), or with ``Pes.host`` a potential that lives on the
    host: the caller's own plugin, called back by the library for whole chains."""

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

    @classmethod
    def host(cls, ndim, natom, potforce, hessian=None):
        """A plugin with no device twin (mcmod_waterdimer.f90:12-103's role, in Python):
        potforce(x) with x of shape (ndof, nbead) returns (grad (ndof, nbead), energy
        (nbead,)) with the plugin's own V - V0; hessian(x) (optional) returns
        (nbead, ndof, ndof) with [b, i, k] = d grad_k / d x_i.  Without it the library
        differences potforce (step 1e-4).  The plain path and the ring routines run with
        it; the batched engine refuses it."""
        self = cls.__new__(cls)
        L.init()
        self.spec = None
        self.ndim, self.natom, self.ndof = ndim, natom, ndim * natom
        D = self.ndof

        def _pf(user, nbead, x, grad, energy):
            try:
                xa = np.ctypeslib.as_array(x, shape=(D * nbead,)).reshape(D, nbead)
                g, e = potforce(xa.copy())
                np.ctypeslib.as_array(grad, shape=(D * nbead,))[:] = np.asarray(g, float).reshape(-1)
                np.ctypeslib.as_array(energy, shape=(nbead,))[:] = np.asarray(e, float).reshape(-1)
                return 0
            except Exception:
                return 1

        def _hs(user, nbead, x, hess):
            try:
                xa = np.ctypeslib.as_array(x, shape=(D * nbead,)).reshape(D, nbead)
                h = np.asarray(hessian(xa.copy()), float).reshape(-1)
                np.ctypeslib.as_array(hess, shape=(nbead * D * D,))[:] = h
                return 0
            except Exception:
                return 1

        self._cb = (L.HOST_POTFORCE(_pf), L.HOST_HESSIAN(_hs) if hessian else L.HOST_HESSIAN())
        h = C.c_void_p()
        L.check(L.lib().rp_pes_create_host(ndim, natom, self._cb[0], self._cb[1], None, C.byref(h)))
        self.h = h
        self.V0 = 0.0
        return self

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
        if x.ndim == 1:  # a column of a chain is a strided view: copy it to contiguous memory
            return np.ascontiguousarray(x).reshape(self.ndof, 1), 1
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
