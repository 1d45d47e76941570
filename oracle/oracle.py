"""Python face of the CPU oracle -- TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg may
import this module, and only as the checker or the timed CPU baseline.  The
product (the HIP library and the ``rptunnel`` host package) never touches it.

Arrays use the reference's layout: a chain ``x(n, ndim, natom)`` stored
column-major, i.e. a numpy array of shape ``(ndof, n)`` with
``dof = atom*ndim + dim``.  Each wrapper names the reference routine it restates.
The pieces written here in numpy (``detj_eigs``, ``ti_quadrature``,
``splitting_*``) restate driver arithmetic that needs no C.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "_build", "liboracle.so")
_LIB_FAST = os.path.join(_HERE, "_build", "liboracle_fast.so")  # the CPU baseline's timed build
_lib = None
_lib_fast = None

PES_QUARTIC, PES_COUPLED2D, PES_EVB, PES_WATER, PES_HARMONIC = 1, 2, 3, 4, 5
OR_PI = 3.14159265358979  # instantonmod.f90:4

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib(fast=False):
    """The parity build, or (fast=True) the -O3 build the CPU baseline times."""
    global _lib, _lib_fast
    if fast:
        if _lib_fast is None:
            if not os.path.exists(_LIB_FAST):
                build()
            _lib_fast = C.CDLL(_LIB_FAST)
            _declare(_lib_fast)
        return _lib_fast
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        _lib = C.CDLL(_LIB)
        _declare(_lib)
    return _lib


class OrPes(C.Structure):
    _fields_ = [("kind", C.c_int), ("ndim", C.c_int), ("natom", C.c_int), ("ndof", C.c_int),
                ("v0", C.c_double), ("dp", _dp), ("ndp", C.c_int), ("ip", _ip), ("nip", C.c_int)]


class OrRng(C.Structure):
    _fields_ = [("normals", _dp), ("nnormals", C.c_int64), ("cursor", C.c_int64),
                ("kicks", C.POINTER(C.c_int32)), ("nkicks", C.c_int64), ("kcursor", C.c_int64),
                ("seed", C.c_uint64), ("stream", C.c_uint64), ("error", C.c_int)]


class OrNm(C.Structure):
    _fields_ = [("n", C.c_int), ("ndim", C.c_int), ("natom", C.c_int), ("ndof", C.c_int),
                ("betan", C.c_double), ("tau", C.c_double), ("dt", C.c_double),
                ("gamma", C.c_double), ("cayley", C.c_int), ("mass", _dp), ("lam", _dp),
                ("beadmass", _dp), ("T", _dp), ("beadvec", _dp), ("c1", _dp), ("c2", _dp),
                ("work", _dp)]


class OrPath(C.Structure):
    _fields_ = [("npath", C.c_int), ("ndof", C.c_int), ("lampath", _dp), ("path", _dp),
                ("y2", _dp)]


class OrTi(C.Structure):
    _fields_ = [("nmc", C.c_int64), ("imin", C.c_int64), ("restartnmc", C.c_int64),
                ("dhdrlimit", C.c_double), ("dbdl", _dp), ("xi", C.c_double),
                ("path", C.POINTER(OrPath)), ("nreinit", C.c_int64), ("dhdr_sum", C.c_double)]


class OrChain(C.Structure):
    _fields_ = [("n", C.c_int), ("ndim", C.c_int), ("natom", C.c_int), ("ndof", C.c_int),
                ("betan", C.c_double), ("mass", _dp), ("fixedends", C.c_int)]


def _declare(L):
    P = C.POINTER
    sig = {
        "or_pes_V": (C.c_double, [P(OrPes), _dp]),
        "or_pes_Vprime": (None, [P(OrPes), _dp, _dp]),
        "or_pes_potforce": (C.c_double, [P(OrPes), _dp, _dp]),
        "or_pes_Vdoubleprime": (None, [P(OrPes), _dp, _dp]),
        "or_pes_param_grad": (C.c_int, [P(OrPes), _dp, _dp]),
        "or_philox_normal": (C.c_double, [C.c_uint64, C.c_uint64, C.c_int64]),
        "or_philox_poisson": (C.c_int32, [C.c_uint64, C.c_uint64, C.c_int64, C.c_double]),
        "or_philox4x32_10": (None, [P(C.c_uint32), P(C.c_uint32), P(C.c_uint32)]),
        "or_nm_init": (C.c_int, [P(OrNm), C.c_int, C.c_int, C.c_int, _dp, C.c_double, C.c_double,
                                 C.c_double, C.c_double, C.c_int, _dp, _dp]),
        "or_nm_free": (None, [P(OrNm)]),
        "or_nm_forward": (None, [P(OrNm), _dp, _dp, C.c_int]),
        "or_nm_backward": (None, [P(OrNm), _dp, _dp, C.c_int]),
        "or_step_v": (C.c_int, [P(OrNm), P(OrPes), C.c_double, _dp, _dp]),
        "or_step_nm": (C.c_int, [P(OrNm), C.c_double, _dp, _dp]),
        "or_step_langevin": (C.c_int, [P(OrNm), P(OrRng), _dp]),
        "or_time_step_pile": (C.c_int, [P(OrNm), P(OrPes), P(OrRng), _dp, _dp]),
        "or_time_step_nm": (C.c_int, [P(OrNm), P(OrPes), _dp, _dp]),
        "or_andersen_resample": (None, [P(OrNm), P(OrRng), _dp]),
        "or_spline": (None, [_dp, _dp, C.c_int, C.c_double, C.c_double, _dp]),
        "or_splint": (C.c_double, [_dp, _dp, _dp, C.c_int, C.c_double]),
        "or_splin_grad": (C.c_double, [_dp, _dp, _dp, C.c_int, C.c_double]),
        "or_path_init": (C.c_int, [P(OrPath), C.c_int, C.c_int, _dp, _dp]),
        "or_path_free": (None, [P(OrPath)]),
        "or_init_path": (None, [P(OrNm), P(OrPath), P(OrRng), C.c_double, _dp, _dp]),
        "or_propagate_pile": (C.c_int, [P(OrNm), P(OrPes), P(OrRng), P(OrTi), _dp, _dp, _dp]),
        "or_propagate_andersen": (C.c_int, [P(OrNm), P(OrPes), P(OrRng), P(OrTi), C.c_int64,
                                            _dp, _dp, _dp]),
        "or_gauleg": (None, [C.c_double, C.c_double, _dp, _dp, C.c_int]),
        "or_UM": (C.c_double, [P(OrChain), P(OrPes), _dp, _dp, _dp]),
        "or_UMprime": (None, [P(OrChain), P(OrPes), _dp, _dp, _dp, _dp]),
        "or_UMforceenergy": (C.c_double, [P(OrChain), P(OrPes), _dp, _dp, _dp, _dp]),
        "or_UMhessian": (None, [P(OrChain), P(OrPes), _dp, C.c_int, _dp]),
        "or_use_blas_dsymv": (None, [C.c_void_p]),
        "or_bench_pile": (C.c_double, [P(OrPes), C.c_int, C.c_int, C.c_int, _dp, C.c_double,
                                       C.c_double, C.c_double, C.c_double, _dp, _dp, _dp, _dp,
                                       C.c_int, C.c_int64, C.c_int, C.c_uint64, _dp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args


def _d(a):
    return a.ctypes.data_as(_dp)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class Pes:
    """A potential in the oracle; ``dp``/``ip`` follow the layouts in oracle.c."""

    def __init__(self, kind, ndim, natom, dp, ip=None, v0=0.0):
        self.kind, self.ndim, self.natom = kind, ndim, natom
        self.dp = _f64(dp)
        self.ip = np.ascontiguousarray(ip if ip is not None else [0], dtype=np.int32)
        self.s = OrPes(kind, ndim, natom, ndim * natom, v0, _d(self.dp), self.dp.size,
                       self.ip.ctypes.data_as(_ip), self.ip.size)

    @property
    def ndof(self):
        return self.ndim * self.natom

    def set_v0(self, v0):
        self.s.v0 = v0

    def V(self, x):
        x = _f64(x)
        return lib().or_pes_V(C.byref(self.s), _d(x))

    def Vprime(self, x):
        x = _f64(x)
        g = np.zeros(self.ndof)
        lib().or_pes_Vprime(C.byref(self.s), _d(x), _d(g))
        return g

    def potforce(self, x):
        x = _f64(x)
        g = np.zeros(self.ndof)
        e = lib().or_pes_potforce(C.byref(self.s), _d(x), _d(g))
        return g, e

    def Vdoubleprime(self, x):
        """hess[i, k] = d grad_k / d x_i (reference hess(i,j,k,l) flattened)."""
        x = _f64(x)
        h = np.zeros(self.ndof * self.ndof)
        lib().or_pes_Vdoubleprime(C.byref(self.s), _d(x), _d(h))
        return h.reshape(self.ndof, self.ndof, order="F")

    def param_grad(self, x):
        x = _f64(x)
        out = np.zeros(4)
        k = lib().or_pes_param_grad(C.byref(self.s), _d(x), _d(out))
        return out[:k]


class Rng:
    """Sequential normal stream: injected array (parity mode) or Philox."""

    def __init__(self, normals=None, kicks=None, seed=0, stream=0):
        self._n = None if normals is None else _f64(normals)
        self._k = None if kicks is None else np.ascontiguousarray(kicks, dtype=np.int32)
        self.s = OrRng()
        self.s.seed, self.s.stream = seed, stream
        if self._n is not None:
            self.s.normals, self.s.nnormals = _d(self._n), self._n.size
        if self._k is not None:
            self.s.kicks = self._k.ctypes.data_as(C.POINTER(C.c_int32))
            self.s.nkicks = self._k.size

    @property
    def cursor(self):
        return self.s.cursor


def philox_normals(seed, stream, start, count):
    L = lib()
    return np.array([L.or_philox_normal(seed, stream, start + i) for i in range(count)])


def philox_poisson(seed, stream, index, mean):
    return lib().or_philox_poisson(seed, stream, index, mean)


class NM:
    """init_nm state (verletmodule.f90:306-338) plus the Langevin c1/c2 table (:380-386)."""

    def __init__(self, n, ndim, natom, mass, betan, tau, dt=1.0, gamma=0.0, cayley=False,
                 a=None, b=None):
        D = ndim * natom
        self.n, self.ndim, self.natom, self.ndof = n, ndim, natom, D
        self.mass = _f64(mass)
        self.a = _f64(np.zeros(D) if a is None else a)
        self.b = _f64(np.zeros(D) if b is None else b)
        self.s = OrNm()
        rc = lib().or_nm_init(C.byref(self.s), n, ndim, natom, _d(self.mass), betan, tau, dt,
                              gamma, int(cayley), _d(self.a), _d(self.b))
        if rc:
            raise FloatingPointError("NaN in init_nm")

    def __del__(self):
        try:
            lib().or_nm_free(C.byref(self.s))
        except Exception:
            pass

    def _arr(self, ptr, count):
        return np.ctypeslib.as_array(ptr, shape=(count,)).copy()

    @property
    def T(self):
        return self._arr(self.s.T, self.n * self.n).reshape(self.n, self.n)

    @property
    def lam(self):
        return self._arr(self.s.lam, self.n)

    @property
    def beadmass(self):
        """shape (n, natom): beadmass(j, i) of the reference at [i, j]."""
        return self._arr(self.s.beadmass, self.n * self.natom).reshape(self.n, self.natom)

    @property
    def beadvec(self):
        """shape (ndof, n)."""
        return self._arr(self.s.beadvec, self.n * self.ndof).reshape(self.ndof, self.n)

    @property
    def c1(self):
        return self._arr(self.s.c1, self.n * self.natom).reshape(self.n, self.natom)

    def forward(self, x, dof=-1):
        x = _f64(x)
        q = np.zeros(self.n)
        lib().or_nm_forward(C.byref(self.s), _d(x), _d(q), dof)
        return q

    def backward(self, q, dof=-1):
        q = _f64(q)
        x = np.zeros(self.n)
        lib().or_nm_backward(C.byref(self.s), _d(q), _d(x), dof)
        return x

    def _chk(self, rc):
        if rc == 3:
            raise FloatingPointError("NaN in propagation")
        if rc == 4:
            raise IndexError("injected noise stream exhausted")
        return rc

    def step_v(self, pes, h, x, p):
        x, p = _f64(x).copy(), _f64(p).copy()
        self._chk(lib().or_step_v(C.byref(self.s), C.byref(pes.s), h, _d(x), _d(p)))
        return x, p

    def step_nm(self, h, x, p):
        x, p = _f64(x).copy(), _f64(p).copy()
        self._chk(lib().or_step_nm(C.byref(self.s), h, _d(x), _d(p)))
        return x, p

    def step_langevin(self, rng, p):
        p = _f64(p).copy()
        self._chk(lib().or_step_langevin(C.byref(self.s), C.byref(rng.s), _d(p)))
        return p

    def time_step_pile(self, pes, rng, x, p):
        x, p = _f64(x).copy(), _f64(p).copy()
        self._chk(lib().or_time_step_pile(C.byref(self.s), C.byref(pes.s), C.byref(rng.s),
                                          _d(x), _d(p)))
        return x, p

    def time_step_nm(self, pes, x, p):
        x, p = _f64(x).copy(), _f64(p).copy()
        self._chk(lib().or_time_step_nm(C.byref(self.s), C.byref(pes.s), _d(x), _d(p)))
        return x, p

    def andersen_resample(self, rng):
        p = np.zeros((self.ndof, self.n))
        lib().or_andersen_resample(C.byref(self.s), C.byref(rng.s), _d(p))
        return p


class Path:
    """Natural-spline path (instantonmod.f90:397-442, 895-1035 minus file I/O)."""

    def __init__(self, lampath, path):
        self.lampath = _f64(lampath)
        self.path = _f64(path)  # (ndof, npath)
        self.s = OrPath()
        lib().or_path_init(C.byref(self.s), self.lampath.size, self.path.shape[0],
                           _d(self.lampath), _d(self.path))

    def __del__(self):
        try:
            lib().or_path_free(C.byref(self.s))
        except Exception:
            pass

    @property
    def y2(self):
        n = self.lampath.size * self.path.shape[0]
        return np.ctypeslib.as_array(self.s.y2, shape=(n,)).copy().reshape(self.path.shape)

    def splint(self, xi):
        L, npth = lib(), self.lampath.size
        y2 = self.y2
        return np.array([L.or_splint(_d(self.lampath), _d(np.ascontiguousarray(self.path[d])),
                                     _d(np.ascontiguousarray(y2[d])), npth, xi)
                         for d in range(self.path.shape[0])])

    def splin_grad(self, xi):
        L, npth = lib(), self.lampath.size
        y2 = self.y2
        return np.array([L.or_splin_grad(_d(self.lampath), _d(np.ascontiguousarray(self.path[d])),
                                         _d(np.ascontiguousarray(y2[d])), npth, xi)
                         for d in range(self.path.shape[0])])


def init_path(nm, path, rng, xi):
    x = np.zeros((nm.ndof, nm.n))
    p = np.zeros((nm.ndof, nm.n))
    lib().or_init_path(C.byref(nm.s), C.byref(path.s), C.byref(rng.s), xi, _d(x), _d(p))
    return x, p


def propagate_pile(nm, pes, rng, x, p, nmc, dbdl, imin=0, dhdrlimit=-1.0, xi=0.0, path=None):
    """propagate_pimd_pile (verletmodule.f90:372-416) -> (x, p, dHdr, nreinit)."""
    x, p = _f64(x).copy(), _f64(p).copy()
    dbdl = _f64(dbdl)
    ti = OrTi(nmc, imin, 0, dhdrlimit, _d(dbdl), xi, C.pointer(path.s) if path else None, 0, 0.0)
    out = C.c_double(0.0)
    nm._chk(lib().or_propagate_pile(C.byref(nm.s), C.byref(pes.s), C.byref(rng.s), C.byref(ti),
                                    _d(x), _d(p), C.byref(out)))
    return x, p, out.value, ti.nreinit


def propagate_andersen(nm, pes, rng, x, p, nmc, dbdl, noutput, imin=0):
    """propagate_pimd_nm (verletmodule.f90:190-250) -> (x, p, dHdr)."""
    x, p = _f64(x).copy(), _f64(p).copy()
    dbdl = _f64(dbdl)
    ti = OrTi(nmc, imin, 0, -1.0, _d(dbdl), 0.0, None, 0, 0.0)
    out = C.c_double(0.0)
    nm._chk(lib().or_propagate_andersen(C.byref(nm.s), C.byref(pes.s), C.byref(rng.s),
                                        C.byref(ti), noutput, _d(x), _d(p), C.byref(out)))
    return x, p, out.value


def gauleg(x1, x2, n):
    x, w = np.zeros(n), np.zeros(n)
    lib().or_gauleg(x1, x2, _d(x), _d(w), n)
    return x, w


class Chain:
    """instantonmod's module state for UM & co: n, masses, betan, fixedends."""

    def __init__(self, n, ndim, natom, mass, betan, fixedends=True):
        self.mass = _f64(mass)
        self.n, self.ndim, self.natom, self.ndof = n, ndim, natom, ndim * natom
        self.s = OrChain(n, ndim, natom, ndim * natom, betan, _d(self.mass), int(fixedends))

    def _ab(self, a, b):
        z = np.zeros(self.ndof)
        return _f64(z if a is None else a), _f64(z if b is None else b)

    def UM(self, pes, x, a=None, b=None):
        a, b = self._ab(a, b)
        return lib().or_UM(C.byref(self.s), C.byref(pes.s), _d(_f64(x)), _d(a), _d(b))

    def UMprime(self, pes, x, a=None, b=None):
        a, b = self._ab(a, b)
        g = np.zeros((self.ndof, self.n))
        lib().or_UMprime(C.byref(self.s), C.byref(pes.s), _d(_f64(x)), _d(a), _d(b), _d(g))
        return g

    def UMforceenergy(self, pes, x, a=None, b=None):
        a, b = self._ab(a, b)
        g = np.zeros((self.ndof, self.n))
        e = lib().or_UMforceenergy(C.byref(self.s), C.byref(pes.s), _d(_f64(x)), _d(a), _d(b),
                                   _d(g))
        return g, e

    def UMhessian(self, pes, x, singlewell=False):
        """LAPACK lower band storage, shape (ndof+1, n*ndof)."""
        band = np.zeros((self.ndof + 1) * self.n * self.ndof)
        lib().or_UMhessian(C.byref(self.s), C.byref(pes.s), _d(_f64(x)), int(singlewell), _d(band))
        return band.reshape(self.ndof + 1, self.n * self.ndof)


def band_to_dense(band):
    """Symmetric dense matrix from LAPACK lower band storage AB(1+i-j, j) = A(i, j)."""
    kd1, n = band.shape
    A = np.zeros((n, n))
    for r in range(kd1):
        j = np.arange(0, n - r)
        A[j + r, j] = band[r, : n - r]
        A[j, j + r] = band[r, : n - r]
    return A


def detj_eigs(band):
    """detJ (instantonmod.f90:782-827): all eigenvalues of the band matrix, ascending
    (DSBEVD's order).  numpy's eigvalsh stands in for the LAPACK call."""
    return np.linalg.eigvalsh(band_to_dense(band))


def lndetj0(eigs):
    """rpi_ser.f90:227-236: sum of ln(eta) over every positive eigenvalue."""
    e = np.asarray(eigs)
    return float(np.sum(np.log(e[e > 0.0])))


def lndetj(eigs):
    """rpi_ser.f90:355-365: skip eta_1, then sum ln(eta) over the positive ones."""
    e = np.asarray(eigs)[1:]
    return float(np.sum(np.log(e[e > 0.0])))


def rpi_splitting(betan, skink, lndetj_val, lndetj0_val):
    """rpi_ser.f90:368-381 -> (phi, theta, delta)."""
    phi = np.exp(0.5 * (lndetj_val - lndetj0_val))
    theta = betan * np.exp(-skink) * np.sqrt(skink / (2.0 * OR_PI)) / phi
    return phi, theta, 2.0 * theta / betan


def ti_quadrature(integrands, weights, nrep, betan):
    """pimd_par.f90:397-424.  integrands: (nintegral*nrep,), lambda-major.
    Returns dict(deltaA, sigmaA, q_ratio, q_err)."""
    I = np.asarray(integrands, dtype=np.float64).reshape(len(weights), nrep)
    mean = I.sum(axis=1) / nrep
    var = (I ** 2).sum(axis=1) / nrep - mean ** 2
    answer = 0.0
    sigmaA = 0.0
    for i in range(len(weights)):
        if mean[i] == mean[i]:
            answer += weights[i] * mean[i]
        sigmaA += var[i] * weights[i] ** 2
    finalI = np.exp(-answer * betan)
    return {"deltaA": answer, "sigmaA": sigmaA, "q_ratio": finalI,
            "q_err": betan * np.sqrt(sigmaA / nrep) * finalI}


def splitting_two_beta(Ia, Ib, beta1, beta2):
    """multiwell_waterdimer.py:6-10."""
    return 2.0 * (np.arctanh(Ib) - np.arctanh(Ia)) / (beta2 - beta1)


def bench_pile(pes, n, ndim, natom, mass, betan, tau, dt, gamma, a, b, dbdl, x0, ntraj, nsteps,
               nthreads, seed, fast=False):
    """Wall seconds for ntraj PILE trajectories (the CPU baseline) and their TI means."""
    mass, a, b, dbdl, x0 = (_f64(v) for v in (mass, a, b, dbdl, x0))
    out = np.zeros(ntraj)
    t = lib(fast).or_bench_pile(C.byref(pes.s), n, ndim, natom, _d(mass), betan, tau, dt, gamma,
                                _d(a), _d(b), _d(dbdl), _d(x0), ntraj, nsteps, nthreads, seed,
                                _d(out))
    return t, out


def use_blas_dsymv(fn_address, fast=True):
    """Route the transforms' mat-vec through a BLAS dsymv_ at fn_address (None: the loop)."""
    lib(fast).or_use_blas_dsymv(fn_address)
