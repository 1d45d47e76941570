"""Potentials and synthetic workloads of the ring-polymer hot path.

None of these potentials exists in the reference; each is a new ``mcmod_mass``
plugin written against the interface of mcmod_waterdimer.f90:12-103 (bohr,
Hartree, ``V - V0``).  The parameter arrays ``dp``/``ip`` use the layouts listed in
csrc/rp_pes.h.  Everything random is drawn from a seeded numpy generator, so the
same spec is rebuilt identically on any machine.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

PES_QUARTIC, PES_COUPLED2D, PES_EVB, PES_WATER, PES_HARMONIC = 1, 2, 3, 4, 5

AMU = 1822.888486  # electron masses per dalton
M_H, M_C, M_N, M_O = 1.00782503 * AMU, 12.0 * AMU, 14.0030740 * AMU, 15.9949146 * AMU


@dataclass
class PesSpec:
    name: str
    kind: int
    ndim: int
    natom: int
    dp: np.ndarray
    ip: np.ndarray | None
    mass: np.ndarray            # natom
    well1: np.ndarray           # ndof, dof = atom*ndim + dim
    well2: np.ndarray
    extra: dict = field(default_factory=dict)

    @property
    def ndof(self):
        return self.ndim * self.natom


def quartic(V0=0.005, a=0.7, mass=1836.0, ndim=1, natom=1):
    """sum_d V0((x_d/a)^2 - 1)^2 -- mcmod_1d.f90:20's form with Vheight=V0, x0=a."""
    D = ndim * natom
    return PesSpec("quartic", PES_QUARTIC, ndim, natom, np.array([V0, a], float), None,
                   np.full(natom, float(mass)), np.full(D, -a), np.full(D, a))


def coupled2d(V0=0.005, a=0.7, w=0.01, c=0.3, m=1836.0):
    """V0(x^2/a^2-1)^2 + m w^2 (y + c x^2)^2 / 2; minima at (+-a, -c a^2)."""
    return PesSpec("coupled2d", PES_COUPLED2D, 2, 1, np.array([V0, a, m, w, c], float), None,
                   np.array([m]), np.array([-a, -c * a * a]), np.array([a, -c * a * a]))


def harmonic(k, c, mass, ndim=1):
    k = np.asarray(k, float)
    c = np.asarray(c, float)
    natom = k.size // ndim
    return PesSpec("harmonic", PES_HARMONIC, ndim, natom, np.concatenate([k, c]), None,
                   np.asarray(mass, float), c.copy(), c.copy())


def evb(natom=9, seed=27, kappa=2.0e-3, lam_lo=1.0e-3, lam_hi=0.3, cubic=1.0e-4, swap=0.6):
    """Two-state EVB double well on 3*natom dof.

    VA = y.Lam.y/2 + sum c y^3 with y = Q^T (x - xa), Lam log-spaced on
    [lam_lo, lam_hi], Q orthogonal from the QR of a seeded Gaussian matrix, |c| <=
    ``cubic``; VB(x) = VA(Pi x) with Pi swapping atoms 0 and 1 (both hydrogen, so
    the two wells are mass-symmetric); V = (VA+VB)/2 - sqrt((VA-VB)^2/4 + kappa^2).
    Atoms 0 and 1 sit ``swap`` bohr apart, which sets the barrier.
    """
    rng = np.random.default_rng(seed)
    ndim = 3
    D = ndim * natom
    lam = np.logspace(np.log10(lam_lo), np.log10(lam_hi), D)
    G = rng.standard_normal((D, D))
    Q, R = np.linalg.qr(G)
    Q = Q * np.sign(np.diag(R))
    cub = rng.uniform(-cubic, cubic, D)
    pos = rng.standard_normal((natom, 3)) * 1.5
    mid = 0.5 * (pos[0] + pos[1])
    u = rng.standard_normal(3)
    u /= np.linalg.norm(u)
    pos[0] = mid - 0.5 * swap * u
    pos[1] = mid + 0.5 * swap * u
    xa = pos.reshape(-1)
    perm = np.arange(natom, dtype=np.int32)
    perm[0], perm[1] = 1, 0
    heavy = [M_C, M_N, M_O]
    mass = np.array([M_H, M_H] + [heavy[i % 3] for i in range(natom - 2)])
    dp = np.concatenate([[kappa], Q.reshape(-1, order="F"), lam, cub, xa])
    well2 = pos[perm].reshape(-1)
    return PesSpec("evb", PES_EVB, ndim, natom, dp, perm, mass, xa.copy(), well2,
                   {"Q": Q, "lam": lam, "cub": cub, "perm": perm})


def water_cluster(nmol=6, seed=6, De=0.185, alpha=1.21, re=1.81, kbend=0.16,
                  theta0=np.deg2rad(104.52), eps=2.4e-4, sigma=5.954, qH=0.41,
                  roo=5.2, jitter=0.05):
    """Flexible water cluster: Morse O-H (De, alpha, re), harmonic HOH bend (kbend,
    theta0), O-O Lennard-Jones (eps, sigma), point charges qH and qO = -2 qH between
    molecules.  Start: a puckered ring of nmol molecules with O-O = ``roo`` bohr and
    seeded Gaussian jitter.  well2 exchanges the two H of molecule 0 (a C2 flip, an
    exact symmetry of V); ``extra['path']`` rotates that molecule's H pair about its
    bisector through 0..180 degrees so the path never makes the two H coincide."""
    rng = np.random.default_rng(seed)
    R = roo / (2.0 * np.sin(np.pi / nmol)) if nmol > 2 else roo / 2.0
    atoms = []
    for m in range(nmol):
        ph = 2.0 * np.pi * m / nmol
        O = np.array([R * np.cos(ph), R * np.sin(ph), 0.45 * (-1) ** m])
        nxt = 2.0 * np.pi * ((m + 1) % nmol) / nmol
        On = np.array([R * np.cos(nxt), R * np.sin(nxt), 0.45 * (-1) ** (m + 1)])
        e1 = On - O if nmol > 1 else np.array([1.0, 0.0, 0.0])  # a monomer has no neighbour
        e1 /= np.linalg.norm(e1)
        zax = np.array([0.0, 0.0, 1.0])
        e2 = np.cross(zax, e1)
        e2 /= np.linalg.norm(e2)
        e3 = np.cross(e1, e2)
        H1 = O + re * e1
        H2 = O + re * (np.cos(theta0) * e1 + np.sin(theta0) * e3)
        atoms += [O, H1, H2]
    pos = np.array(atoms) + rng.standard_normal((3 * nmol, 3)) * jitter
    mass = np.array([M_O, M_H, M_H] * nmol)
    w1 = pos.reshape(-1).copy()
    pos2 = pos.copy()
    pos2[[1, 2]] = pos[[2, 1]]
    w2 = pos2.reshape(-1)
    # rotation of molecule 0's H pair about its HOH bisector
    O, H1, H2 = pos[0], pos[1], pos[2]
    bis = (H1 - O) / np.linalg.norm(H1 - O) + (H2 - O) / np.linalg.norm(H2 - O)
    bis /= np.linalg.norm(bis)
    frames = []
    for ang in np.linspace(0.0, np.pi, 9):
        c, s = np.cos(ang), np.sin(ang)
        K = np.array([[0, -bis[2], bis[1]], [bis[2], 0, -bis[0]], [-bis[1], bis[0], 0]])
        Rm = np.eye(3) + s * K + (1 - c) * K @ K
        f = pos.copy()
        f[1] = O + Rm @ (H1 - O)
        f[2] = O + Rm @ (H2 - O)
        frames.append(f.reshape(-1))
    frames[-1] = w2.copy()
    dp = np.array([De, alpha, re, kbend, theta0, eps, sigma, qH], float)
    return PesSpec("water", PES_WATER, 3, 3 * nmol, dp, None, mass, w1, w2,
                   {"path": np.array(frames)})


def linear_path(well1, well2, npath=2):
    """Straight path well1 -> well2 sampled at npath points: (lampath, path[ndof, npath]).
    lampath is the normalised arc length, as read_path builds it
    (instantonmod.f90:926-936)."""
    t = np.linspace(0.0, 1.0, npath)
    path = np.outer(well1, 1.0 - t) + np.outer(well2, t)
    return arc_lampath(path), path


def arc_lampath(path):
    """Normalised arc length of a path stored (ndof, npath)."""
    from .inputs import arc_lampath as frames_lampath

    return frames_lampath(np.asarray(path, dtype=np.float64).T)
