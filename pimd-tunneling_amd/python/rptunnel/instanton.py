"""Host mirror of instantonmod (instantonmod.f90) on the device ring kernels.

``Ring`` holds a batch of open chains of uneven length.  Methods follow the reference
routines: ``UM`` / ``UMprime`` / ``UMforceenergy`` (:17-151), ``UMhessian`` (:155-217),
``instanton`` (:679-777), ``detJ``-style log-determinants (:782-827 with the sums of
rpi_ser.f90:227-236, 355-365).  A chain's coordinates use the reference layout
x(n, ndim, natom) column-major, i.e. numpy shape (ndof, n).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from .pes import Pes


class Ring:
    def __init__(self, pes: Pes, mass, n, betan, a=None, b=None, fixedends=True):
        self.pes = pes
        self.ndof = pes.ndof
        self.n = [int(v) for v in np.atleast_1d(n)]
        self.nchain = len(self.n)
        self.betan = L.f64(np.broadcast_to(np.asarray(betan, dtype=np.float64), (self.nchain,)))
        D = self.ndof
        za = np.zeros((self.nchain, D))
        self.a = L.f64(za if a is None else np.broadcast_to(a, (self.nchain, D)))
        self.b = L.f64(za if b is None else np.broadcast_to(b, (self.nchain, D)))
        self.mass = L.f64(mass)
        narr = np.ascontiguousarray(self.n, dtype=np.int32)
        h = C.c_void_p()
        L.check(L.lib().rp_ring_create(pes.h, L.ptr(self.mass), self.nchain, L.iptr(narr),
                                       L.ptr(self.betan), L.ptr(self.a), L.ptr(self.b),
                                       int(fixedends), C.byref(h)))
        self.h = h
        self.off = np.concatenate([[0], np.cumsum([k * D for k in self.n])])

    def close(self):
        if getattr(self, "h", None):
            L.lib().rp_ring_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- chains <-> flat buffer
    def pack(self, chains):
        return L.f64(np.concatenate([np.asarray(c, dtype=np.float64).reshape(-1) for c in chains]))

    def unpack(self, flat):
        return [flat[self.off[c]:self.off[c + 1]].reshape(self.ndof, self.n[c])
                for c in range(self.nchain)]

    def set_x(self, chains):
        L.check(L.lib().rp_ring_set_x(self.h, L.ptr(self.pack(chains))))

    def get_x(self):
        x = np.zeros(self.off[-1])
        L.check(L.lib().rp_ring_get_x(self.h, L.ptr(x)))
        return self.unpack(x)

    # -- chains already in device memory: flat float64 buffers of length off[-1] on the
    # library's GPU, passed by address, complete before the call (rp_ring_*_device)
    def set_x_device(self, xptr):
        L.check(L.lib().rp_ring_set_x_device(self.h, C.c_void_p(xptr)))

    def um_device(self, gptr=None):
        """U_M per chain (host); the adjoint gradient into the device buffer at gptr."""
        e = np.zeros(self.nchain)
        L.check(L.lib().rp_ring_um_device(self.h, L.ptr(e), C.c_void_p(gptr) if gptr else None))
        return e

    def hvp_device(self, vptr, hvptr):
        L.check(L.lib().rp_ring_hvp_device(self.h, C.c_void_p(vptr), C.c_void_p(hvptr)))

    # -- reference routines
    def UMforceenergy(self, chains=None):
        """-> (gradients per chain, U_M per chain)"""
        if chains is not None:
            self.set_x(chains)
        e = np.zeros(self.nchain)
        g = np.zeros(self.off[-1])
        L.check(L.lib().rp_ring_um(self.h, L.ptr(e), L.ptr(g)))
        return self.unpack(g), e

    def UM(self, chains=None):
        if chains is not None:
            self.set_x(chains)
        e = np.zeros(self.nchain)
        L.check(L.lib().rp_ring_um(self.h, L.ptr(e), None))
        return e

    def UMprime(self, chains=None):
        return self.UMforceenergy(chains)[0]

    def hvp(self, v_chains):
        v = self.pack(v_chains)
        hv = np.zeros_like(v)
        L.check(L.lib().rp_ring_hvp(self.h, L.ptr(v), L.ptr(hv)))
        return self.unpack(hv)

    def UMhessian(self, chain=0, singlewell=False):
        """LAPACK lower band storage AB(ndof+1, n*ndof), with the reference's coupling;
        the library fills it column-major, as the reference's answer(ndof+1, n*ndof)."""
        D, n = self.ndof, self.n[chain]
        band = np.zeros((D + 1) * n * D)
        L.check(L.lib().rp_ring_hessian_band(self.h, chain, int(singlewell), L.ptr(band)))
        return band.reshape(n * D, D + 1).T

    def logdet(self, singlewell=False, skip_lowest=True, want_eta=None):
        """-> (sum of ln(eta) over the eigenvalues the reference keeps, nneg, eta_1).
        eta_1 (the lowest eigenvalue: with none negative, inverse iteration shifted below it
        on a bracket of inertia counts; else the one nearest zero by inverse iteration and
        any other negative ones by bisection on the inertia) is computed when skip_lowest
        needs it or want_eta asks for it; otherwise it is returned as NaN."""
        if want_eta is None:
            want_eta = skip_lowest
        ld = np.zeros(self.nchain)
        nn = np.zeros(self.nchain, dtype=np.int32)
        eta = np.full(self.nchain, np.nan)
        L.check(L.lib().rp_ring_logdet(self.h, int(singlewell), int(skip_lowest), L.ptr(ld),
                                       L.iptr(nn), L.ptr(eta) if want_eta else None))
        return ld, nn, eta

    def lowest_eigs(self, k, singlewell=False):
        """The k lowest eigenvalues of every chain's band (detJ's ascending spectrum,
        instantonmod.f90:782-827, head only): shape (nchain, k)."""
        out = np.zeros(self.nchain * k)
        L.check(L.lib().rp_ring_lowest_eigs(self.h, int(singlewell), int(k), L.ptr(out)))
        return out.reshape(self.nchain, k)

    def set_preconditioner(self, sigma):
        """rp_ring_set_preconditioner: L-BFGS with gamma P^-1 as its initial inverse Hessian,
        P = the springs' tridiagonal + sigma I per dof column; sigma <= 0 switches it off."""
        L.check(L.lib().rp_ring_set_preconditioner(self.h, float(sigma)))

    def instanton(self, chains, pgtol=1e-5, factr=1e6, m=8, maxiter=0):
        """L-BFGS on every chain at once -> (optimised chains, U_M, iterations, status)."""
        x = self.pack(chains)
        act = np.zeros(self.nchain)
        it = np.zeros(self.nchain, dtype=np.int32)
        st = np.zeros(self.nchain, dtype=np.int32)
        L.check(L.lib().rp_instanton(self.h, pgtol, factr, m, maxiter, L.ptr(x), L.ptr(act),
                                     L.iptr(it), L.iptr(st)))
        return self.unpack(x), act, it, st

    def param_grad(self):
        out = np.zeros((self.nchain, 2))
        npar = C.c_int(0)
        L.check(L.lib().rp_ring_param_grad(self.h, L.ptr(out), C.byref(npar)))
        return out[:, : npar.value]
