"""The batched engine: many trajectories (TI nodes x replicas x bead counts) advanced
together by the fused persistent step kernel, with propagate_pimd_pile /
propagate_pimd_nm semantics per polymer (verletmodule.f90:190-250, 372-416)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from .pes import Pes

LANGEVIN, ANDERSEN = 2, 1


class Batch:
    """polymers: sequence of dicts with keys n, betan, xi, a, b, dbdl, stream."""

    def __init__(self, pes: Pes, mass, polymers, tau, dt, gamma=0.0, thermostat=LANGEVIN,
                 noutput=1, dhdrlimit=-1.0, cayley=False, seed=0):
        self.pes = pes
        self.ndof = pes.ndof
        self.mass = L.f64(mass)
        self.n = [int(p["n"]) for p in polymers]
        self._keep = []
        descs = (L.PolymerDesc * len(polymers))()
        for k, p in enumerate(polymers):
            a, b, d = (L.f64(p[key]) for key in ("a", "b", "dbdl"))
            self._keep += [a, b, d]
            descs[k] = L.PolymerDesc(int(p["n"]), float(p["betan"]), float(p.get("xi", 0.0)),
                                     L.ptr(a), L.ptr(b), L.ptr(d), int(p.get("stream", k)))
        h = C.c_void_p()
        L.check(L.lib().rp_batch_create(pes.h, L.ptr(self.mass), len(polymers), descs, tau, dt,
                                        gamma, thermostat, noutput, dhdrlimit, int(cayley), seed,
                                        C.byref(h)))
        self.h = h
        self.npoly = len(polymers)

    def close(self):
        if getattr(self, "h", None):
            L.lib().rp_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_path(self, lampath, path):
        lampath, path = L.f64(lampath), L.f64(path)
        L.check(L.lib().rp_batch_set_path(self.h, lampath.size, L.ptr(lampath), L.ptr(path)))

    def init_path(self, noise=None):
        nz = None if noise is None else L.f64(noise)
        L.check(L.lib().rp_batch_init_path(self.h, L.ptr(nz)))

    def set_state(self, poly, x, p):
        L.check(L.lib().rp_batch_set_state(self.h, poly, L.ptr(L.f64(x)), L.ptr(L.f64(p))))

    def get_state(self, poly):
        n = self.n[poly]
        x = np.zeros((self.ndof, n))
        p = np.zeros((self.ndof, n))
        L.check(L.lib().rp_batch_get_state(self.h, poly, L.ptr(x), L.ptr(p)))
        return x, p

    def set_kicks(self, kicks):
        if kicks is None:
            L.check(L.lib().rp_batch_set_kicks(self.h, None, 0))
            return
        kk = np.ascontiguousarray(kicks, dtype=np.int32)
        L.check(L.lib().rp_batch_set_kicks(self.h, kk.ctypes.data_as(L.i32p), kk.shape[1]))

    def run(self, nsteps, imin=0, noise=None):
        """noise: None (Philox) or array (npoly, nnoise_per_poly) of injected draws."""
        if noise is None:
            L.check(L.lib().rp_batch_run(self.h, nsteps, imin, None, 0))
        else:
            nz = L.f64(noise)
            L.check(L.lib().rp_batch_run(self.h, nsteps, imin, L.ptr(nz), nz.shape[1]))

    def ti_readout(self):
        s = np.zeros(self.npoly)
        c = np.zeros(self.npoly, dtype=np.int64)
        r = np.zeros(self.npoly, dtype=np.int64)
        L.check(L.lib().rp_batch_ti_readout(self.h, L.ptr(s), c.ctypes.data_as(L.i64p),
                                            r.ctypes.data_as(L.i64p)))
        return s, c, r

    def cursors(self):
        c = np.zeros(self.npoly, dtype=np.int64)
        L.check(L.lib().rp_batch_get_cursor(self.h, c.ctypes.data_as(L.i64p)))
        return c

    def kick_cursors(self):
        c = np.zeros(self.npoly, dtype=np.int64)
        L.check(L.lib().rp_batch_get_kick_cursor(self.h, c.ctypes.data_as(L.i64p)))
        return c

    def set_cursors(self, cursor=None, kick_cursor=None):
        """Stream positions of every polymer (None keeps them), e.g. from restart files."""
        c = None if cursor is None else np.ascontiguousarray(cursor, dtype=np.int64)
        k = None if kick_cursor is None else np.ascontiguousarray(kick_cursor, dtype=np.int64)
        for a in (c, k):
            if a is not None and a.shape != (self.npoly,):
                raise ValueError(f"set_cursors: need {self.npoly} cursors, got {a.shape}")
        L.check(L.lib().rp_batch_set_cursor(
            self.h, None if c is None else c.ctypes.data_as(L.i64p),
            None if k is None else k.ctypes.data_as(L.i64p)))

    def families(self):
        nf = C.c_int(0)
        n = np.zeros(64, dtype=np.int32)
        cnt = np.zeros(64, dtype=np.int32)
        L.check(L.lib().rp_batch_families(self.h, C.byref(nf), L.iptr(n), L.iptr(cnt)))
        return list(zip(n[: nf.value].tolist(), cnt[: nf.value].tolist()))

    def family_shapes(self):
        """Per family: (n, waves per workgroup, MFMA units per wave, folded transforms)."""
        out = []
        for f, (n, _) in enumerate(self.families()):
            w, u, fo = C.c_int(0), C.c_int(0), C.c_int(0)
            L.check(L.lib().rp_batch_family_shape(self.h, f, C.byref(w), C.byref(u), C.byref(fo)))
            out.append((n, w.value, u.value, bool(fo.value)))
        return out

    def set_launch(self, steps_per_launch=0, waves=0):
        L.check(L.lib().rp_batch_set_launch(self.h, steps_per_launch, waves))

    def set_graphs(self, enable=True):
        """Replay step blocks as captured HIP graphs (default) or launch kernels directly."""
        L.check(L.lib().rp_batch_set_graphs(self.h, int(bool(enable))))

    def autotune(self, trial_steps=64):
        bs, bw = C.c_int(0), C.c_int(0)
        L.check(L.lib().rp_batch_autotune(self.h, trial_steps, C.byref(bs), C.byref(bw)))
        return bs.value, bw.value

    def last_kernel_ms(self):
        nf = len(self.families())
        ms = np.zeros(max(nf, 1))
        nl = C.c_int64(0)
        L.check(L.lib().rp_batch_last_kernel_ms(self.h, L.ptr(ms), C.byref(nl)))
        return ms[:nf], nl.value
