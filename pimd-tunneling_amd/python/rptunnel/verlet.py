"""Host mirror of the reference's ``verletint`` module (verletmodule.f90) on the
plain per-trajectory GPU path.

``Chain`` holds one trajectory's state on the device.  Its methods are named after
the reference routines they replace and take the same arguments; arrays use the
reference layout x(n, ndim, natom) column-major, i.e. numpy shape ``(ndof, n)``.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from .pes import Pes


class Chain:
    """init_nm (verletmodule.f90:306-338) + the per-trajectory integrator state."""

    def __init__(self, pes: Pes, n, mass, betan, tau, dt=1.0, gamma=0.0, a=None, b=None,
                 cayley=False):
        self.pes = pes
        self.n, self.ndim, self.natom, self.ndof = n, pes.ndim, pes.natom, pes.ndof
        self.mass = L.f64(mass)
        self.betan, self.tau, self.dt, self.gamma = betan, tau, dt, gamma
        a = L.f64(np.zeros(self.ndof) if a is None else a)
        b = L.f64(np.zeros(self.ndof) if b is None else b)
        h = C.c_void_p()
        L.check(L.lib().rp_nm_setup(pes.h, n, self.ndim, self.natom, L.ptr(self.mass), betan,
                                    tau, dt, gamma, int(cayley), L.ptr(a), L.ptr(b), C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            L.lib().rp_chain_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ------------------------------------------------------------ state
    def init_nm(self, a, b):
        """Re-run init_nm's beadvec for new fixed ends (pimd_par.f90:326)."""
        L.check(L.lib().rp_chain_set_ends(self.h, L.ptr(L.f64(a)), L.ptr(L.f64(b))))

    def set_state(self, x, p):
        L.check(L.lib().rp_chain_set_state(self.h, L.ptr(L.f64(x)), L.ptr(L.f64(p))))

    def get_state(self):
        x = np.zeros((self.ndof, self.n))
        p = np.zeros((self.ndof, self.n))
        L.check(L.lib().rp_chain_get_state(self.h, L.ptr(x), L.ptr(p)))
        return x, p

    def nm_arrays(self):
        """(transmatrix, lam, beadmass (n, natom), beadvec (ndof, n))."""
        n = self.n
        T = np.zeros((n, n))
        lam = np.zeros(n)
        bm = np.zeros((n, self.natom))
        bv = np.zeros((self.ndof, n))
        L.check(L.lib().rp_chain_get_nm(self.h, L.ptr(T), L.ptr(lam), L.ptr(bm), L.ptr(bv)))
        return T, lam, bm, bv

    def set_rng(self, seed, stream, cursor=0):
        L.check(L.lib().rp_chain_set_rng(self.h, seed, stream, cursor))

    @property
    def cursor(self):
        c = C.c_int64(0)
        L.check(L.lib().rp_chain_get_cursor(self.h, C.byref(c)))
        return c.value

    # ------------------------------------------------------------ transforms
    def _transform(self, direction, vecs, shift):
        vecs = L.f64(np.atleast_2d(vecs))
        nc = vecs.shape[0]
        sh = np.ascontiguousarray(np.broadcast_to(np.asarray(shift, dtype=np.int32), (nc,)))
        out = np.zeros_like(vecs)
        L.check(L.lib().rp_nm_transform(self.h, direction, nc, L.iptr(sh), L.ptr(vecs),
                                        L.ptr(out)))
        return out

    def nmtransform_forward(self, xprop, bead=0):
        """verletmodule.f90:254-266; bead is the reference's 1-based dof (0 = no shift)."""
        return self._transform(0, xprop, np.asarray(bead) - 1)

    def nmtransform_backward(self, qprop, bead=0):
        """verletmodule.f90:270-286."""
        return self._transform(1, qprop, np.asarray(bead) - 1)

    # ------------------------------------------------------------ integrator
    def step_v(self, time):
        L.check(L.lib().rp_step_v(self.h, time))

    def step_nm(self, time):
        L.check(L.lib().rp_step_nm(self.h, time))

    def step_langevin(self, noise=None):
        L.check(L.lib().rp_step_langevin(self.h, L.ptr(None if noise is None else L.f64(noise))))

    def time_step_pile(self, noise=None):
        L.check(L.lib().rp_time_step_pile(self.h, L.ptr(None if noise is None else L.f64(noise))))

    def time_step_nm(self):
        L.check(L.lib().rp_time_step_nm(self.h))

    def andersen_resample(self, noise=None):
        L.check(L.lib().rp_andersen_resample(self.h, L.ptr(None if noise is None else L.f64(noise))))

    def set_path(self, lampath, path):
        lampath, path = L.f64(lampath), L.f64(path)
        L.check(L.lib().rp_chain_set_path(self.h, lampath.size, L.ptr(lampath), L.ptr(path)))

    def init_path(self, xi, noise=None):
        L.check(L.lib().rp_init_path(self.h, xi, L.ptr(None if noise is None else L.f64(noise))))

    def propagate_pimd_pile(self, nmc, dbdl, imin=0, dhdrlimit=-1.0, xi=0.0, noise=None):
        """verletmodule.f90:372-416 -> (dHdr, nreinit)."""
        out = C.c_double(0.0)
        nre = C.c_int64(0)
        nz = None if noise is None else L.f64(noise)
        L.check(L.lib().rp_propagate_pile(self.h, nmc, imin, dhdrlimit, L.ptr(L.f64(dbdl)), xi,
                                          L.ptr(nz), 0 if nz is None else nz.size,
                                          C.byref(out), C.byref(nre)))
        return out.value, nre.value

    def propagate_pimd_nm(self, nmc, dbdl, noutput, imin=0, noise=None, kicks=None):
        """verletmodule.f90:190-250 (Andersen) -> dHdr."""
        out = C.c_double(0.0)
        nz = None if noise is None else L.f64(noise)
        kk = None if kicks is None else np.ascontiguousarray(kicks, dtype=np.int32)
        L.check(L.lib().rp_propagate_nm(
            self.h, nmc, imin, noutput, L.ptr(L.f64(dbdl)), L.ptr(nz),
            0 if nz is None else nz.size,
            None if kk is None else kk.ctypes.data_as(L.i32p), 0 if kk is None else kk.size,
            C.byref(out)))
        return out.value
